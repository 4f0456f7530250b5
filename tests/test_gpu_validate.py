"""Ghia validation of the GPU backend end to end: the `simple` CLI (backend = gpu) runs the
reference's validation cases to convergence, writes its output files, and validate_ghia (the
restatement of postprocess/validate_ghia.py) checks the centrelines against Ghia et al. (1982)
within the reference's CTest limits (CMakeLists.txt:139-150), and against the reference's
published GPU results where they exist: 3718 and 5919 outer iterations with rms_u / rms_v
0.01698 / 0.00825 and 0.01257 / 0.01093 at 80^2 and 128^2, Re 1000, multicolour + red-black
(docs/RESULTS.md:329). The 128^2 run writes its fields as .npy (field_format = npy).
"""
import json
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

import validate_ghia as vg  # noqa: E402

CASES = [
    # (config, extra keys, expected iterations or None, published (rms_u, rms_v) or None)
    ("re1000_80", {}, 3718, ("0.01698", "0.00825")),
    ("re1000_128", {"field_format": "npy"}, 5919, ("0.01257", "0.01093")),
    ("re400_80", {}, None, None),
    ("re400_80", {"pressure_solver": "mg"}, None, None),
    ("re1000_80", {"pressure_solver": "mg"}, None, None),
]


@pytest.mark.parametrize("name,extra,iters,published", CASES)
def test_ghia_validation_gpu(tmp_path, name, extra, iters, published):
    text = open(os.path.join(GOLDEN, "ghia", name + ".cfg")).read()
    out = tmp_path / name
    lines = [ln for ln in text.splitlines()
             if ln.split("=")[0].strip() not in set(extra) | {"output_dir"}]
    lines += [f"{k} = {v}" for k, v in extra.items()] + [f"output_dir = {out}"]
    cfg = tmp_path / (name + ".cfg")
    cfg.write_text("\n".join(lines) + "\n")
    r = subprocess.run([os.path.join(ROOT, "build", "simple"), str(cfg)], capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    info = json.loads((out / "run_info.json").read_text())
    ext = ".npy" if extra.get("field_format") == "npy" else ".csv"
    assert (out / ("u" + ext)).exists() and (out / ("ux_centre" + ext)).exists()
    row = vg.evaluate(str(out))
    print(f"{name} {extra}: " + "  ".join(vg._fmt(row)) + f"  ({info['seconds']:.2f} s)",
          flush=True)
    assert info["converged"]
    if iters is not None:
        assert row[3] == iters
    if published is not None:
        assert (f"{row[5]:.5f}", f"{row[7]:.5f}") == published
    max_u, max_v = vg.LIMITS[row[1]]
    assert row[5] <= max_u and row[7] <= max_v, row


def test_bench_runner_on_gpu(tmp_path):
    """The fixed-work runner (restating bench/run_bench.py) on the GPU CLI: one row per run with
    every phase of timing.csv, the solver time of run_info.json and no convergence at tolerance 0.
    The phase times are measured on sampled eager iterations, so their sum is compared with the
    solver time only loosely."""
    import csv

    import run_bench as rb
    out = tmp_path / "gpu.csv"
    rows = rb.run([64, 96], 30, 1, [1], os.path.join(ROOT, "build", "simple"), "gpu",
                  {"mg_precision": "mixed"}, out)
    with open(out) as fh:
        got = list(csv.DictReader(fh))
    assert [int(r["grid"]) for r in got] == [64, 96]
    for r in got:
        print({k: r[k] for k in ("grid", "iters", "seconds_per_iter", "solve_p", "cpu")},
              flush=True)
        assert int(r["iters"]) == 30 and r["converged"] == "False"
        assert float(r["seconds_per_iter"]) > 0 and float(r["solve_p"]) > 0
        assert r["cpu"] and r["compiler"] != "unknown"
        phases = sum(float(r[p]) for p in ("assemble_u", "sweep_u", "assemble_p", "solve_p",
                                           "correct"))
        assert phases < 3 * float(r["total_seconds"])
    assert len(rows) == 2
