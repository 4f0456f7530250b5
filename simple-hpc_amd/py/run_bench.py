"""Fixed-work benchmark runner for the `backend = gpu` CLI: one CSV row per run, in the columns of
the reference's runner (bench/run_bench.py:98-185).

What it restates:
 - the config of a run: the 80 x 80 Re 1000 cavity template of the reference's runner
   (configs/cavity_re1000_80.cfg, its keys and values below) with QUICK, diffusion order 4,
   Re 1000 and error tolerance 0 forced (bench/run_bench.py:25-30), then N_cells_x = N_cells_y =
   grid, maxit = iters, output_dir = a temporary directory and any --set KEY=VALUE
   (:33-55); the GPU path's own keys (backend = gpu, momentum_solver = gs_multicolour,
   pressure_solver = mg) are the defaults here, and --set overrides them too;
 - one run: the CLI on that config, then timing.csv's per-phase seconds and run_info.json's
   solver seconds and iteration count (:98-112);
 - the row: tag, grid, threads, repeat, iters, total_seconds, seconds_per_iter,
   cell_updates_per_s (grid^2 x iters / seconds), one column per phase, cpu, compiler, git_hash,
   converged, cg_iters_mean (:160-185). `threads` is the value passed (the GPU path does not use
   it); `cpu` holds run_info's `device` (the GPU) when there is one, else the host CPU's name;
 - the order: repeat-major, the thread list rotated every repeat (:150-156).

Usage: python run_bench.py [--grids 256 512 1024] [--iters 200] [--repeats 3] [--threads 1]
           [--exe build/simple] [--tag gpu] [--set KEY=VALUE ...] [--out results.csv]
           [--write-config DIR]

--write-config DIR writes the config of each grid's run to DIR/g<grid>.cfg (output_dir
DIR/g<grid>) and runs nothing: the same run under a profiler is then `simple DIR/g<grid>.cfg`.
"""
import argparse
import csv
import datetime
import json
import os
import platform
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent.parent

# configs/cavity_re1000_80.cfg of the reference, key by key.
TEMPLATE = [
    ("N_cells_x", "80"), ("N_cells_y", "80"), ("Re", "1000"), ("maxit", "15000"),
    ("convection_scheme", "QUICK"), ("diffusion_order", "4"), ("N_halo_x", "2"),
    ("N_halo_y", "2"), ("maxit_inner_u", "4"), ("maxit_inner_v", "4"), ("maxit_inner_p", "8"),
    ("error_tolerance", "0.0001"), ("relaxation_factor_u", "0.3"), ("relaxation_factor_v", "0.3"),
    ("relaxation_factor_p", "0.2"), ("Hybrid_switch", "False"), ("backend", "serial"),
    ("momentum_solver", "gs_lex"), ("pressure_solver", "gs_lex"),
    ("output_dir", "results/cavity_re1000_80"),
]
OVERRIDES = {"convection_scheme": "QUICK", "diffusion_order": "4", "Re": "1000",
             "error_tolerance": "0"}
GPU_PATH = {"backend": "gpu", "momentum_solver": "gs_multicolour", "pressure_solver": "mg"}
FIXED = ["tag", "grid", "threads", "repeat", "iters", "total_seconds", "seconds_per_iter",
         "cell_updates_per_s"]
TRAILER = ["cpu", "compiler", "git_hash", "converged", "cg_iters_mean"]


def make_config(grid, iters, out_dir, extra=None):
    """Config text: the template with the benchmark keys, the GPU path and `extra` applied."""
    values = dict(OVERRIDES)
    values.update(GPU_PATH)
    values.update(extra or {})
    values["N_cells_x"] = str(grid)
    values["N_cells_y"] = str(grid)
    values["maxit"] = str(iters)
    values["output_dir"] = str(out_dir)
    lines, seen = [], set()
    for key, val in TEMPLATE:
        lines.append(f"{key} = {values.get(key, val)}")
        seen.add(key)
    lines += [f"{k} = {v}" for k, v in values.items() if k not in seen]
    return "\n".join(lines) + "\n"


def read_timing(path):
    """[(phase, seconds)] of timing.csv, in file order, `total` included."""
    with open(path, newline="") as fh:
        return [(r["phase"], float(r["seconds"])) for r in csv.DictReader(fh)]


def read_run_info(path):
    if not Path(path).exists():
        return {}
    with open(path) as fh:
        return json.load(fh)


def host_cpu():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine() or "unknown"


def run_one(exe, grid, threads, iters, extra=None):
    """One CLI run: (phases, run_info). Raises with the tail of the output if it fails."""
    with tempfile.TemporaryDirectory(prefix="bench_") as tmp:
        out = Path(tmp) / "out"
        cfg = Path(tmp) / "bench.cfg"
        cfg.write_text(make_config(grid, iters, out, extra))
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        p = subprocess.run([str(exe), str(cfg)], env=env, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError(f"run failed (grid={grid}, threads={threads}):\n"
                               + p.stdout[-2000:] + p.stderr[-2000:])
        return read_timing(out / "timing.csv"), read_run_info(out / "run_info.json")


def record(tag, grid, threads, rep, iters_asked, phases, info, phase_names, cpu):
    times = dict(phases)
    total = float(info.get("seconds", times["total"]))  # solver time, output excluded
    iters = int(info.get("iterations", iters_asked))
    row = {"tag": tag, "grid": grid, "threads": threads, "repeat": rep, "iters": iters,
           "total_seconds": f"{total:.9f}", "seconds_per_iter": f"{total / iters:.9f}",
           "cell_updates_per_s": f"{grid * grid * iters / total:.6e}"}
    row.update({p: f"{times[p]:.9f}" for p in phase_names})
    row.update({"cpu": info.get("device") or cpu, "compiler": info.get("compiler", "unknown"),
                "git_hash": info.get("git_hash", "unknown"),
                "converged": info.get("converged", ""),
                "cg_iters_mean": info.get("cg_iterations_mean", "")})
    return row


def run(grids, iters, repeats, threads, exe, tag, extra, out_path, runner=run_one, log=None):
    records, phase_names, cpu = [], None, host_cpu()
    for rep in range(1, repeats + 1):
        k = (rep - 1) % len(threads)
        order = threads[k:] + threads[:k]
        for grid in grids:
            for t in order:
                phases, info = runner(exe, grid, t, iters, extra)
                if phase_names is None:
                    phase_names = [p for p, _ in phases if p not in ("total", "unaccounted")]
                r = record(tag, grid, t, rep, iters, phases, info, phase_names, cpu)
                records.append(r)
                if log:
                    log(f"grid={grid} threads={t} repeat={rep} iters={r['iters']} "
                        f"seconds_per_iter={float(r['seconds_per_iter']):.6f}")
    with open(out_path, "w", newline="") as fh:
        w = csv.DictWriter(fh, fieldnames=FIXED + (phase_names or []) + TRAILER)
        w.writeheader()
        w.writerows(records)
    return records


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--grids", type=int, nargs="+", default=[256, 512, 1024])
    ap.add_argument("--iters", type=int, default=200)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--threads", type=int, nargs="+", default=[1])
    ap.add_argument("--exe", default=str(ROOT / "build" / "simple"))
    ap.add_argument("--tag", default="gpu")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE")
    ap.add_argument("--out", default=None, help="CSV path (default: <tag>_<stamp>.csv here)")
    ap.add_argument("--write-config", default=None, metavar="DIR",
                    help="write DIR/g<grid>.cfg for each grid and run nothing")
    a = ap.parse_args(argv)
    extra = {}
    for item in a.set:
        if "=" not in item:
            sys.exit(f"--set expects KEY=VALUE, got: {item}")
        k, v = item.split("=", 1)
        extra[k.strip()] = v.strip()
    if a.write_config:
        d = Path(a.write_config).resolve()
        d.mkdir(parents=True, exist_ok=True)
        for grid in a.grids:
            path = d / f"g{grid}.cfg"
            path.write_text(make_config(grid, a.iters, d / f"g{grid}", extra))
            print(path)
        return 0
    exe = Path(a.exe).resolve()
    if not exe.exists():
        sys.exit(f"executable not found: {exe}")
    out = a.out or f"{a.tag}_{datetime.datetime.now().strftime('%Y%m%d-%H%M')}.csv"
    run(a.grids, a.iters, a.repeats, a.threads, exe, a.tag, extra, out,
        log=lambda m: print(m, file=sys.stderr, flush=True))
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
