"""The reference-side adapter (simple-hpc_amd/integration/solver_sg.cpp) as the reference would use
it: the reference's HipSolver and run_simple_hip implemented over this repository's C ABI,
compiled against the reference's own headers and linked with the reference's own CPU code into
oracle/_ref/adapter_check (oracle/ref.mk, built where the reference sources are).

adapter_check makes the adapter's exact call sequence: HipSolver(p, state) uploads all 28 State
arrays (sg_create + sg_upload), each phase runs through the adapter and on the reference's CPU
functions, and after every phase HipSolver::download fills a State whose 28 arrays must equal the
CPU's bit for bit (the shape of tests/hip/test_hip.cpp:70-153); then run_simple_hip (the
reference's run_outer_loop over sg_outer_iteration) against run_simple, logs and arrays bit for
bit.
"""
import os
import subprocess

import pytest

from conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

CHECK = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
CFG = os.path.join(GOLDEN, "snapshots", "converged_32x40.cfg")


def _run(cfg, *overrides, timeout=300):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref/adapter_check is not built (needs the reference sources)")
    env = dict(os.environ, OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS", "8"))
    r = subprocess.run([CHECK, cfg, *overrides], capture_output=True, text=True,
                       timeout=timeout, env=env)
    lines = [x for x in r.stdout.splitlines() if x.startswith(("PASS", "FAIL", "ALL", "case"))]
    print("\n".join(lines), flush=True)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
    assert "ALL PASSED" in r.stdout
    return r.stdout


@pytest.mark.parametrize("scheme,order,mom,pres", [
    ("QUICK", 4, "gs_multicolour", "gs_redblack"),
    ("UDS", 2, "gs_multicolour", "gs_redblack"),
    ("QUICK", 4, "gs_multicolour", "mg"),
    ("QUICK", 4, "gs_multicolour", "cg"),
    ("QUICK", 2, "jacobi", "jacobi"),
])
def test_adapter_per_phase_and_run(scheme, order, mom, pres):
    _run(CFG, f"convection_scheme={scheme}", f"diffusion_order={order}",
         f"momentum_solver={mom}", f"pressure_solver={pres}", "maxit=200", "error_tolerance=0")


def test_adapter_bench_workload_1024(tmp_path):
    """The adapter on the bench workload's size (1024^2, Re 1000, QUICK / 4th order, multicolour +
    multigrid V(2,2)): per phase, then 12 outer iterations of run_simple_hip."""
    cfg = tmp_path / "bench.cfg"
    cfg.write_text("N_cells_x = 1024\nN_cells_y = 1024\nRe = 1000\nconvection_scheme = QUICK\n"
                   "diffusion_order = 4\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
                   "mg_pre = 2\nmg_post = 2\nerror_tolerance = 0\nmaxit = 12\nbackend = gpu\n"
                   "write_fields = False\n")
    _run(str(cfg), timeout=600)
