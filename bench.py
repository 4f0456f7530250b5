"""Benchmark of the SIMPLE lid-driven-cavity hot path on one MI355X.

Default workload (N=1): the 1024 x 1024 cavity, Re 1000, QUICK convection, 4th-order diffusion,
multicolour Gauss-Seidel momentum (4 sweeps), geometric-multigrid V(2,2) p', zero initial fields,
absolute SIMPLE residual tolerance 1e-5 (the stop rule of the reference's run_outer_loop). A step
is one SIMPLE outer iteration over the whole grid, replayed as a captured hipGraph; the fields stay
resident in HBM, the stop rule runs on the device, and every timed region is bracketed by device
synchronisations.

The metric is time to the 1e-5 residual, plain FP64 against learned + mixed precision:
  plain    the reference's arithmetic, bit-identical to its CPU path;
  learned  mixed-precision multigrid (FP32 V-cycle inside FP64 defect correction) started from the
           warm-start U-Net's prediction of p' (bfloat16 MFMA; simple-hpc_amd/weights/); the
           outer loop's stop rule is the same.
Both run from zero fields to the tolerance in the same process, with the launch shapes of one
autotune. `value` is the learned + mixed time; the plain time, the ratio and the largest
differences of u, v and p between the two converged states are on the same line. Also on the
line:
  - per-iteration times of every form (`--steps` iterations each after `--warmup`): plain,
    mixed (FP32 V-cycle), mixed_mom (also FP32 momentum corrections inside FP64 defect
    correction, momentum_precision = mixed) and learned;
  - `roofline`: the dominant kernel (one u multicolour sweep) timed with hipEvents on the
    solver's stream, credited with its algorithmic bytes per cell (the reference sweep's arrays
    read and written once: 9 coefficients, p and u read, u written, 96 B), against 8 TB/s; the
    HBM traffic per launch from the committed rocprofv3 counter summary; beside it
    (`outer_iteration`) the plain form's whole outer iteration credited the same way (about
    1851 B per cell);
  - `cpu_baseline`: the reference's own OpenMP CPU code (built from its sources into oracle/_ref;
    the oracle restating it where that build is absent) on this host's cores, on a bounded
    sample of the same workload, and its projection to the tolerance.

Other workloads (--workload):
  batch    the 64 uneven cavities of simple-hpc_amd/workloads/cavities_seed0.txt to 1e-5, one
           persistent launch per phase, mixed and FP64 forms;
  large    one 8192 x 8192 cavity, Re 3200: fixed-work outer iterations/s of plain, mixed and
           learned + mixed;
  learned  the warm start's training (data from batched runs, GPU training) and its A/B;
  unet     the network's bfloat16 MFMA kernels alone.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--mode tol|fixed] [--forms ...]
                       [--workload cavity|batch|large|learned|unet] [--knobs ...] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simple-hpc_amd", "py"))

HBM_PEAK_GBS = 8000.0
# The reference sweep's algorithmic bytes per cell: 9 coefficients, p and u read, u written.
SWEEP_BYTES_PER_CELL = 96.0
# A whole FP64 outer iteration with multigrid V(2,2) p', each array of the reference's passes read
# and written once per logical pass, per cell: assemble_u 88, residual_u 88, 4 u sweeps 4 x 96,
# the same for v, assemble_p 80, the p' solve about 579 (level 0: smoothing 4 x 64, residual 64,
# restriction 12, prolongation 18, operator build 42, times 4/3 for the coarse levels; mean
# projections 48, zero fill 8), correct 72.
ITERATION_BYTES_PER_CELL = 88 + 88 + 4 * 96 + 88 + 88 + 4 * 96 + 80 + (4 * 64 + 64 + 12 + 18 + 42) * 4 / 3 + 48 + 8 + 72
WEIGHTS = os.path.join(ROOT, "simple-hpc_amd", "weights", "unet_warm_start.bin")
# HBM bytes per launch of the dominant kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE runs,
# calibrated on a copy kernel (recorded with the commands that made them).
PROFILE = os.path.join(ROOT, "profiles", "k_mom_sweep_1024sq_re1000.md")
FORMS = ("plain", "mixed", "mixed_mom", "learned")


def recorded_traffic(n):
    if n != 1024 or not os.path.exists(PROFILE):
        return None
    with open(PROFILE) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch ="):
                return float(line.split("=")[1])
    return None


def config_text(n, form, re=1000, maxit=10000000):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = {re}\nconvection_scheme = QUICK\n"
            f"diffusion_order = 4\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_pre = 2\nmg_post = 2\nerror_tolerance = 1e-5\nmaxit = {maxit}\nbackend = gpu\n"
            f"write_fields = False\ncheck_every = 500\n")
    if form in ("mixed", "mixed_mom", "learned"):
        text += "mg_precision = mixed\n"
    if form == "mixed_mom":
        text += "momentum_precision = mixed\n"
    if form == "learned":
        text += "pressure_warm_start = learned\n"
    return text


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


class Form:
    """A solver of one form, with the warm-start network attached for `learned`."""

    def __init__(self, sg, n, form, tune, re=1000, level=None):
        self.p, _ = sg.params_from_config(config_text(n, form, re))
        self.s = sg.Solver(self.p)
        self.net = None
        if tune:
            sg.set_tunings(self.s, tune)
        if form == "learned":
            import warmstart as ws
            self.net = sg.UNet(ws.load(WEIGHTS))
            # -1: the library picks the coarsest multigrid level whose grid is still at least
            # 64 x 64, the size of the coarse grids the network was trained on
            sg.attach_warm_start(self.s, self.net, -1 if level is None else level)
            self.level = sg.warm_level(self.s)

    def close(self):
        self.s.close()
        if self.net is not None:
            self.net.close()


def autotune(sg, n, knobs):
    """Launch shapes {form: knobs}: the FP64 momentum sweeps and FP64 multigrid tuned on the plain
    form, the FP32 multigrid and the FP32 correction sweeps on the mixed-momentum form (the mixed
    forms take the FP64 sweep tiles of the plain one: their momentum sweeps are the same)."""
    if knobs:  # given launch shapes "tile,threads,mid,lds[,staged[,first[,dc]]]" (no autotuner)
        t = tuple(int(x) for x in knobs.split(","))
        return {f: t for f in FORMS}
    tunes = {}
    for form in ("plain", "mixed_mom"):
        f = Form(sg, n, form, None)
        f.s.advance(3)
        rep = f.s.autotune(reps=10)
        for line in rep.strip().splitlines():
            log(f"{n}^2 {form} autotune: {line}")
        tunes[form] = sg.tuning_of(f.s)
        f.close()
    keys = list(sg.TUNE_KEYS)
    mixed = list(tunes["mixed_mom"])
    for k in (sg.TUNE_MOM_TILE, sg.TUNE_MOM_TILE_FIRST):
        mixed[keys.index(k)] = tunes["plain"][keys.index(k)]
    return {"plain": tunes["plain"], "mixed": tuple(mixed), "mixed_mom": tuple(mixed),
            "learned": tuple(mixed)}


def fixed_work(sg, n, form, tune, steps, warmup, level=None):
    f = Form(sg, n, form, tune, level=level)
    f.s.advance(warmup)
    t0 = time.perf_counter()
    done, lg, _ = f.s.advance(steps)
    t = time.perf_counter() - t0
    out = {"form": form, "iterations": done, "seconds": t, "ms_per_iter": 1e3 * t / done,
           "last_residuals": [float(x) for x in lg[-1]]}
    if form == "learned":
        out["guesses_kept_discarded"] = list(sg.warm_stats(f.s))
        out["warm_level"] = f.level
    f.close()
    return out


def to_tolerance(sg, n, form, tune, chunk=5000, level=None):
    f = Form(sg, n, form, tune, level=level)
    log(f"{n}^2 {form}: to tolerance")
    # Every outer iteration from zero fields is timed, the graph captures (a few ms) included.
    t0 = time.perf_counter()
    total, conv, last = 0, False, None
    next_report = t0 + 20
    while not conv and total < f.p.maxit:
        k, lg, conv = f.s.advance(chunk)
        total += k
        last = lg[-1] if k else last
        if time.perf_counter() > next_report:
            log(f"  {form}: {total} iterations, residuals {last}")
            next_report += 20
        if k == 0:
            break
    t = time.perf_counter() - t0
    out = {"form": form, "iterations": total, "seconds": t,
           "ms_per_iter": 1e3 * t / max(total, 1), "converged": bool(conv),
           "last_residuals": [float(x) for x in last]}
    if form == "learned":
        out["guesses_kept_discarded"] = list(sg.warm_stats(f.s))
        out["warm_level"] = f.level
    return f, out


def field_errors(p, a, b):
    import numpy as np
    H = p.H
    du = float(np.abs(a.get("u") - b.get("u")).max())
    dv = float(np.abs(a.get("v") - b.get("v")).max())
    pa = a.get("p")[H:H + p.NX, H + 1:H + 1 + p.NY]
    pb = b.get("p")[H:H + p.NX, H + 1:H + 1 + p.NY]
    dp = float(np.abs((pa - pa.mean()) - (pb - pb.mean())).max())
    return {"u": du, "v": dv, "p": dp}


# This backend's config keys that the reference's parser does not know (it rejects unknown keys).
ADDED_KEYS = {"mg_precision", "mg_coarsest", "mg_dc_maxit", "pressure_warm_start",
              "warm_start_weights", "warm_start_level", "check_every", "field_format", "autotune",
              "momentum_precision"}


def reference_config(text):
    """A config text of this backend turned into the reference's: the added keys dropped and
    backend = openmp (the reference's CPU path)."""
    out = []
    for line in text.splitlines():
        key = line.split("=")[0].strip()
        if key in ADDED_KEYS or key == "backend":
            continue
        out.append(line)
    return "\n".join(out + ["backend = openmp"]) + "\n"


class CpuSolver:
    """FP64 outer iterations on the host: the reference's own OpenMP CPU code built from its
    sources into oracle/_ref (kind "reference") or, where that build is absent, the oracle
    restating it (kind "port"). step() runs one outer iteration and returns its seconds."""

    def __init__(self, sg, text, cores):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        if pyoracle.ref_available():
            pyoracle.ref_lib().ref_set_threads(cores)
            self.o = pyoracle.Reference(reference_config(text))
            self.kind, self.what = "reference", "the reference's own CPU code (oracle/_ref)"
        else:
            pyoracle.set_threads(cores)
            self.o = pyoracle.Oracle(sg.params_from_config(text)[0])
            self.kind, self.what = "port", "the oracle restating the reference's OpenMP path"

    def step(self):
        if self.kind == "port":
            return self.o.run_fixed(1)[0]
        t0 = time.perf_counter()
        self.o.outer_iteration()
        return time.perf_counter() - t0


def cpu_baseline(sg, n, seconds=12.0, iterations=None):
    """FP64 outer iterations of the workload at its full size on all host cores for about
    `seconds` (CpuSolver: the reference's own CPU code where its build is present)."""
    cores = os.cpu_count() or 1
    o = CpuSolver(sg, config_text(n, "plain"), cores)
    k, t = 0, 0.0
    while t < seconds and k < 2000:
        t += o.step()
        k += 1
    out = {"value": k / t, "unit": "outer iterations/s", "cores": cores, "kind": o.kind,
           "sample": f"{k} FP64 outer iterations of the {n}^2 workload from zero fields "
                     f"({o.what}, {cores} OpenMP threads), {t:.1f} s"}
    if iterations:
        # the plain paths are bit-identical, so the CPU needs the same iteration count
        out["projected_time_to_tolerance_s"] = iterations * t / k
    return out


def iteration_roofline(n, form, ms):
    """One whole outer iteration of `form` credited with ITERATION_BYTES_PER_CELL, against 8 TB/s."""
    b = ITERATION_BYTES_PER_CELL * n * n
    achieved = b / (ms * 1e-3) / 1e9
    return {"form": form, "algorithmic_bytes_per_cell": ITERATION_BYTES_PER_CELL,
            "algorithmic_bytes": b, "ms": ms, "achieved": achieved, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS}


def run_cavity(sg, a):
    import numpy as np  # noqa: F401
    n = a.n
    tune = autotune(sg, n, a.knobs)
    forms = [f for f in a.forms.split(",") if f]
    level = a.warm_level
    table = [fixed_work(sg, n, f, tune[f], a.steps, a.warmup, level) for f in FORMS]
    for r in table:
        log(f"{n}^2 {r['form']}: {r['ms_per_iter']:.4f} ms per outer iteration "
            f"({a.steps} after {a.warmup})")
    # the dominant kernel, on a solver of the headline form's momentum arithmetic (FP64)
    probe = Form(sg, n, "plain", tune["plain"])
    probe.s.advance(3)
    sec, _, per = probe.s.probe(sg.K_SWEEP_U, reps=200)
    cells = n * (n - 1)
    csec, cbytes, _ = probe.s.probe(sg.K_COPY, reps=200)
    probe.close()
    achieved = SWEEP_BYTES_PER_CELL * cells / sec / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": recorded_traffic(n),
            "kernel": "k_mom_sweep<5> (one u multicolour sweep, all colours, face form)",
            "kernel_us": sec * 1e6, "algorithmic_bytes_per_launch": SWEEP_BYTES_PER_CELL * cells,
            "algorithmic_bytes_per_cell": SWEEP_BYTES_PER_CELL, "cells_per_launch": cells,
            "launches_per_iteration": per,
            "device_copy": {"bytes_per_launch": cbytes, "us": csec * 1e6,
                            "GBs": cbytes / csec / 1e9},
            "outer_iteration": iteration_roofline(
                n, "plain", next(r for r in table if r["form"] == "plain")["ms_per_iter"])}
    out = {"n_gpus": 1, "warmup": a.warmup, "scaling": "weak", "vs_baseline": None,
           "data": "synthetic (lid-driven cavity from zero fields, lid velocity 1)"}
    workload = (f"lid-driven cavity {n}x{n}, Re 1000, QUICK, 4th-order diffusion, gs_multicolour "
                f"momentum (4 sweeps), mg V(2,2) p', tol 1e-5")
    if a.mode == "fixed":
        head = next(r for r in table if r["form"] == forms[0])
        out.update({"metric": f"outer iterations/s, {n}^2 Re1000 ({head['form']})",
                    "value": head["iterations"] / head["seconds"], "unit": "outer iterations/s",
                    "steps": head["iterations"], "ms_per_step": head["ms_per_iter"],
                    "higher_is_better": True, "dtype": "f64",
                    "config": {"workload": workload, "per_iteration": table,
                               "knobs": {k: list(v) for k, v in tune.items()}}})
    else:
        runs, solvers = {}, {}
        for f in forms:
            solvers[f], runs[f] = to_tolerance(sg, n, f, tune[f], level=level)
            log(f"{n}^2 {f}: {runs[f]['iterations']} iterations in {runs[f]['seconds']:.1f} s")
        head = runs[forms[-1]]
        base = runs[forms[0]]
        errs = field_errors(solvers[forms[0]].p, solvers[forms[-1]].s, solvers[forms[0]].s)
        for s in solvers.values():
            s.close()
        out.update({
            "metric": f"time to SIMPLE residual 1e-5, {n}^2 Re1000: {forms[0]} FP64 vs "
                      f"{'mixed+learned' if forms[-1] == 'learned' else forms[-1]}",
            "value": head["seconds"], "unit": "s", "steps": head["iterations"],
            "ms_per_step": head["ms_per_iter"], "higher_is_better": False,
            "dtype": "f64 (plain); f64 + f32 V-cycle + bf16 network (mixed+learned)",
            "config": {"workload": workload, "to_tolerance": runs,
                       f"speedup_{forms[-1]}_over_{forms[0]}": base["seconds"] / head["seconds"],
                       f"max_abs_diff_{forms[-1]}_vs_{forms[0]}": errs,
                       "per_iteration": table,
                       "knobs": {k: list(v) for k, v in tune.items()},
                       "warm_start": {"weights": os.path.relpath(WEIGHTS, ROOT),
                                      "level": solvers[forms[-1]].level
                                      if forms[-1] == "learned" else None}}})
    out["roofline"] = roof
    if not a.no_cpu:
        log("cpu baseline")
        its = out["config"].get("to_tolerance", {}).get("plain", {}).get("iterations")
        out["cpu_baseline"] = cpu_baseline(sg, n, iterations=its)
    return out


def run_batch(sg, steps_cap, knobs=None, cpu=True):
    """The batched uneven-cavity workload to tolerance, mixed and FP64 forms."""
    import numpy as np
    cases = sg.cavities()
    out = {}
    batches = {}
    tune = None
    for form in ("mixed", "fp64"):
        plist = [sg.params_from_config(sg.batch_config(n, re, precision=form))[0]
                 for n, re in cases]
        b = sg.Batch(plist)
        if tune is None:
            if knobs:
                tune = tuple(int(x) for x in knobs.split(","))
            else:
                b.advance(3)
                rep = b.autotune(reps=5)
                log("batch autotune -> " + rep.strip().splitlines()[-1])
                tune = sg.tuning_of(b)
                b.close()
                b = sg.Batch(plist)
        sg.set_tunings(b, tune)
        log(f"batch {form}: {len(cases)} cavities to tolerance "
            f"({b.launches()} launches per batched iteration)")
        b.advance(1)  # graph capture outside the timed region
        t0 = time.perf_counter()
        total_its, st = None, None
        next_report = t0 + 20
        while True:
            its, st, last = b.advance(2000)
            if all(x != 0 for x in st) or max(its) >= steps_cap:
                total_its = its
                break
            if time.perf_counter() > next_report:
                log(f"  {form}: max iterations {max(its)}, running {st.count(0)}")
                next_report += 20
        t = time.perf_counter() - t0
        out[form] = {"seconds": t, "iterations": total_its, "status": st,
                     "cavity_iterations": int(sum(total_its) - len(cases)),
                     "cavity_iterations_per_s": (sum(total_its) - len(cases)) / t,
                     "converged": sum(1 for x in st if x == 1)}
        batches[form] = b
    errs = []
    for k in range(len(cases)):
        cm, c6 = batches["mixed"].cavity(k), batches["fp64"].cavity(k)
        errs.append([float(np.max(np.abs(cm.get(f) - c6.get(f)))) for f in ("u", "v", "p")])
    for b in batches.values():
        b.close()
    e = np.array(errs)
    res = {"cases": [[n, re] for n, re in cases], "forms": out,
           "max_abs_err_mixed_vs_fp64": {"u": float(e[:, 0].max()), "v": float(e[:, 1].max()),
                                         "p": float(e[:, 2].max())},
           "per_cavity_err": errs, "tuning": list(tune)}
    if cpu:
        res["cpu_baseline"] = cpu_batch_baseline(sg, cases, out["fp64"]["iterations"])
    return res


def cpu_batch_baseline(sg, cases, gpu_iters, seconds=15.0):
    """CPU s/iteration per distinct cavity size (all host cores; CpuSolver: the reference's own
    CPU code where its build is present), projected onto the FP64 iteration counts
    (bit-identical paths, so the CPU would need the same counts)."""
    cores = os.cpu_count() or 1
    per = {}
    budget = seconds / len(set(n for n, _ in cases))
    kind = what = None
    for n in sorted(set(n for n, _ in cases)):
        re = [r for m, r in cases if m == n][0]
        o = CpuSolver(sg, sg.batch_config(n, re, precision="fp64"), cores)
        kind, what = o.kind, o.what
        k, t = 0, 0.0
        while t < budget and k < 200:
            t += o.step()
            k += 1
        per[n] = t / k
    total = sum(per[n] * (its - 1) for (n, _), its in zip(cases, gpu_iters))
    work = sum(its - 1 for its in gpu_iters)
    return {"value": work / total, "unit": "cavity outer iterations/s", "cores": cores,
            "kind": kind, "projected_time_to_tolerance_s": total,
            "s_per_iteration": {str(k): v for k, v in per.items()},
            "sample": f"outer iterations per cavity size ({what}, {budget:.1f} s each, "
                      f"{cores} threads), projected onto the FP64 iteration counts"}


def run_large(sg, steps, warmup, knobs=None):
    n, res = 8192, []
    tune = tuple(int(x) for x in knobs.split(",")) if knobs else None
    for form in ("plain", "mixed", "learned"):
        f = Form(sg, n, form, tune, re=3200)
        f.s.advance(warmup)
        log(f"8192^2 {form}: timing {steps} outer iterations")
        t0 = time.perf_counter()
        done, lg, conv = f.s.advance(steps)
        t = time.perf_counter() - t0
        r = {"form": form, "iterations": done, "seconds": t, "ms_per_iter": 1e3 * t / done,
             "last_residuals": [float(x) for x in lg[-1]]}
        if form == "plain":
            sec, bytes_, per = f.s.probe(sg.K_SWEEP_U, reps=20)
            r["sweep_u_us"] = sec * 1e6
            r["sweep_u_GBs_algorithmic"] = SWEEP_BYTES_PER_CELL * n * (n - 1) / sec / 1e9
            r["outer_iteration_roofline"] = iteration_roofline(n, form, r["ms_per_iter"])
        if form == "learned":
            r["guesses_kept_discarded"] = list(sg.warm_stats(f.s))
            r["warm_level"] = f.level
        res.append(r)
        f.close()
    return res


def run_learned(sg, quick=False, full=False):
    """The learned warm start: data from batched runs, GPU training, held-out error, per-solve and
    end-to-end A/B against the mixed-precision solve without it."""
    import numpy as np
    import warmstart as ws
    level = 2
    snaps = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]
    cases = ws.training_cavities(n=8 if quick else 24, seed=1)
    log(f"learned: data from {len(cases)} batched cavities, {len(snaps)} snapshots")
    t0 = time.perf_counter()
    data = ws.generate(cases, snaps[:6] if quick else snaps, level=level, log=log)
    held = ws.generate(ws.training_cavities(n=5, seed=2), [16, 256, 2048], level=level, log=log)
    t_data = time.perf_counter() - t0
    net = sg.UNet(ws.init_params(0))
    t0 = time.perf_counter()
    steps = 300 if quick else 3000
    hist = ws.train(net, data, steps=steps, batch=8, lr=1e-3, log=log, every=500, augment=True)
    t_train = time.perf_counter() - t0
    mse, mse0 = ws.evaluate(net, held)
    log(f"held-out mse {mse:.4f} vs zero prediction {mse0:.4f}")
    sec256, fl256 = net.probe(1, 256, 256, reps=20)
    per_solve, stats = ws.ab_per_solve(net, 256 if quick else 1024, 1000.0,
                                        [10, 100, 1000], level=level)
    log(f"per-solve p' error (zero start, warm start): {per_solve}, guesses {stats}")
    e2e = ws.ab_to_tolerance(net, 128 if quick else 256, 1000.0, level=level, log=log)
    out = {"pairs": sum(len(v[0]) for v in data.values()),
           "coarse_shapes": sorted([list(k) for k in data]), "data_seconds": t_data,
           "train_steps": steps, "train_seconds": t_train,
           "loss_first_last": [float(np.mean(hist[:50])), float(np.mean(hist[-50:]))],
           "heldout_mse": mse, "heldout_mse_zero": mse0,
           "forward_256sq_us": sec256 * 1e6, "forward_256sq_tflops": fl256 / sec256 / 1e12,
           "per_solve_rel_error": [[a, b, c] for a, b, c in per_solve],
           "per_solve_guesses_kept_discarded": list(stats), "to_tolerance": e2e}
    if full:
        out["to_tolerance_1024"] = ws.ab_to_tolerance(net, 1024, 1000.0, level=level, log=log)
    return out


def run_unet(sg, reps=50):
    """The warm-start network's bfloat16 MFMA kernels: one forward pass at the 1024^2 cavity's
    level-2 shape (256 x 256), at the smaller shapes a coarse-level network would see (the
    multigrid tail starts at 64 x 64), and one training step (8 x 64 x 64), timed with hipEvents."""
    import numpy as np
    import warmstart as ws
    net = sg.UNet(ws.init_params(0))
    sec, fl = net.probe(1, 256, 256, reps=reps)
    small = {f"{h}x{h}": net.probe(1, h, h, reps=reps)[0] * 1e6 for h in (128, 64, 32)}
    rng = np.random.default_rng(0)
    x = rng.standard_normal((8, 64, 64, 8)).astype(np.float32)
    x[..., 3:] = 0
    t = rng.standard_normal((8, 64, 64)).astype(np.float32)
    net.train_step(x, t, 1e-4)
    t0 = time.perf_counter()
    for _ in range(reps):
        net.train_step(x, t, 1e-4)
    tstep = (time.perf_counter() - t0) / reps
    _, fl64 = net.probe(8, 64, 64, reps=1)
    net.close()
    return {"forward_us": sec * 1e6, "forward_flops": fl, "forward_tflops": fl / sec / 1e12,
            "forward_us_smaller": small, "train_step_ms": tstep * 1e3, "train_step_flops": 3 * fl64,
            "train_tflops": 3 * fl64 / tstep / 1e12}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000,
                    help="outer iterations of each form's fixed-work timing (per-iteration table)")
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--mode", default="tol", choices=["fixed", "tol"])
    ap.add_argument("--forms", default="plain,learned",
                    help="to-tolerance forms, the baseline first and the headline last")
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="cavity",
                    choices=["cavity", "batch", "large", "learned", "unet"])
    ap.add_argument("--warm-level", type=int, default=None,
                    help="multigrid level the warm-start network runs on (default: the coarsest "
                         "level with at least 64 x 64 cells)")
    ap.add_argument("--quick", action="store_true", help="learned: a small, fast version")
    ap.add_argument("--full", action="store_true", help="learned: also 1024^2 to tolerance")
    ap.add_argument("--knobs", default=None,
                    help="launch shapes 'tile,tail_threads,mid_cells,lds_cells[,staged[,first"
                         "[,dc]]]' instead of autotuning (profiling runs)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU: the multi-GPU decomposition of the cavity is not built "
              "(DESIGN.md describes how it would shard).", file=sys.stderr)
        sys.exit(2)
    import simple_gpu as sg
    if a.workload == "batch":
        r = run_batch(sg, 10000000, a.knobs, cpu=not a.no_cpu)
        m = r["forms"]["mixed"]
        out = {"metric": "time to SIMPLE residual 1e-5, 64 uneven cavities (mixed, batched)",
               "value": m["seconds"], "unit": "s", "n_gpus": 1, "steps": max(m["iterations"]),
               "warmup": 1, "ms_per_step": 1e3 * m["seconds"] / max(m["iterations"]),
               "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
               "dtype": "f64+f32", "data": "synthetic (lid-driven cavities from zero fields)",
               "config": {"workload": "64 square cavities, N uniform in {64,96,128,192,256}, Re "
                                      "log-uniform [100,3200], mt19937_64 seed 0, QUICK, "
                                      "2nd-order diffusion, gs_multicolour, mg V(2,2), tol 1e-5",
                          **r}}
        if "cpu_baseline" in r:
            out["cpu_baseline"] = r.pop("cpu_baseline")
        print(json.dumps(out))
        return
    if a.workload == "unet":
        r = run_unet(sg)
        peak = 2500.0  # MI355X dense bf16 MFMA, TFLOP/s
        out = {"metric": "U-Net forward pass, 256x256 (bf16 MFMA, FP32 accumulation)",
               "value": r["forward_tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": 50,
               "warmup": 1, "ms_per_step": r["forward_us"] / 1e3, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
               "data": "synthetic (random weights and inputs)",
               "config": {"workload": "warm-start U-Net 16-32-64-128, 3x3 convs, one 256x256 "
                                      "image; training step 8x64x64", **r},
               "roofline": {"bound": "mfma", "achieved": r["forward_tflops"], "peak": peak,
                            "unit": "TFLOP/s", "frac": r["forward_tflops"] / peak,
                            "traffic": None}}
        print(json.dumps(out))
        return
    if a.workload == "learned":
        r = run_learned(sg, quick=a.quick, full=a.full)
        e = r["to_tolerance"]
        out = {"metric": "time to SIMPLE residual 1e-5 with the learned warm start (learned + "
                         "mixed vs mixed)",
               "value": e["learned"]["seconds"], "unit": "s", "n_gpus": 1,
               "steps": e["learned"]["iterations"], "warmup": 1,
               "ms_per_step": e["learned"]["ms_per_iter"], "higher_is_better": False,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32+bf16",
               "data": "synthetic (lid-driven cavities from zero fields)",
               "config": {"workload": "U-Net warm start (16-32-64-128, bf16 MFMA) trained on "
                                      "batched-run pairs, A/B against mixed multigrid", **r}}
        print(json.dumps(out))
        return
    if a.workload == "large":
        r = run_large(sg, a.steps if a.steps != 1000 else 50, a.warmup if a.warmup != 50 else 3,
                      a.knobs)
        out = {"metric": "outer iterations/s, 8192^2 Re3200 (plain)",
               "value": r[0]["iterations"] / r[0]["seconds"], "unit": "outer iterations/s",
               "n_gpus": 1, "steps": r[0]["iterations"], "warmup": a.warmup,
               "ms_per_step": r[0]["ms_per_iter"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f64",
               "data": "synthetic (lid-driven cavity from zero fields, lid velocity 1)",
               "config": {"workload": "lid-driven cavity 8192x8192, Re 3200, QUICK, 4th-order "
                                      "diffusion, gs_multicolour, mg V(2,2)", "forms": r}}
        print(json.dumps(out))
        return
    print(json.dumps(run_cavity(sg, a)))


if __name__ == "__main__":
    main()
