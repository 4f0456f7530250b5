"""The learned warm start with its committed weights (simple-hpc_amd/weights/unet_warm_start.bin,
trained by simple-hpc_amd/py/train_warm_start.py): what the trained network gains over the method
without it, and what it does not.

 - Per p' solve: on states of a 256^2 and of the 1024^2 Re 1000 run, one mixed-precision p'
   solve started from the network's guess ends closer to the converged p' than the same solve
   started from zero (the gain), and the guesses are kept.
 - End to end, 256^2 Re 1000 to the 1e-5 residual: plain FP64, mixed, and learned + mixed all
   converge; the learned form's outer-iteration count stays within 1 % of the mixed form's (the
   SIMPLE iteration count is set by the under-relaxation, not by how exactly p' is solved), and
   its converged u, v, p lie within 1e-6 of the FP64 run's (itself bit-identical to the oracle).
   Times are printed: the network's forward pass costs more per iteration than it saves.
 - The stop rule is the reference's: the run stops at the first iteration after the first whose
   largest residual is at most the tolerance (src/solver.cpp:125).
 - The bound: a p' solved ten times more exactly (four V-cycles) leaves the outer-iteration count
   unchanged, so no p'-side method can gain more than the p' solve's share of an iteration
   (below 2x at 1024^2, where the target asks for 3x).
"""
import os

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

import simple_gpu as sg  # noqa: E402
import warmstart as ws  # noqa: E402

WEIGHTS = os.path.join(ROOT, "simple-hpc_amd", "weights", "unet_warm_start.bin")


def _params(n, form, tol=1e-5, check_every=500):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = 1000\nconvection_scheme = QUICK\n"
            f"diffusion_order = 4\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_pre = 2\nmg_post = 2\nerror_tolerance = {tol!r}\nmaxit = 10000000\n"
            f"backend = gpu\nwrite_fields = False\ncheck_every = {check_every}\n")
    if form != "plain":
        text += "mg_precision = mixed\n"
    if form == "learned":
        text += "pressure_warm_start = learned\n"
    return sg.params_from_config(text)[0]


@pytest.fixture(scope="module")
def net():
    assert os.path.exists(WEIGHTS), "the trained weights are committed with the repository"
    u = sg.UNet(ws.load(WEIGHTS))
    yield u
    u.close()


@pytest.mark.parametrize("n,snaps", [(256, [10, 100, 1000, 4000]), (1024, [10, 100, 1000])])
def test_warm_start_gain_per_solve(net, n, snaps):
    """The network on the automatic level (a 64 x 64 coarse grid: level 2 at 256^2, level 4 at
    1024^2) cuts the p' error of one solve to at most a quarter of the zero start's (measured:
    0.05-0.11x at 256^2, 0.12-0.16x at 1024^2)."""
    per, (kept, rej) = ws.ab_per_solve(net, n, 1000.0, snaps)
    for snap, e0, e1 in per:
        print(f"{n}^2 iteration {snap}: p' error after one solve from zero {e0:.3f}, "
              f"from the network's guess {e1:.3f} ({e1 / e0:.2f}x)", flush=True)
    assert kept == len(snaps) and rej == 0
    for snap, e0, e1 in per:
        assert e1 < 0.25 * e0, (n, snap, e0, e1)


def _to_tolerance(p, net=None):
    s = sg.Solver(p)
    if net is not None:
        sg.attach_warm_start(s, net, -1)
    import time
    t0 = time.perf_counter()
    total, conv, logs = 0, False, []
    while not conv and total < 200000:
        k, lg, conv = s.advance(5000)
        total += k
        logs.append(lg)
        if k == 0:
            break
    return s, total, conv, time.perf_counter() - t0, np.concatenate(logs)


@pytest.fixture(scope="module")
def runs_256(net):
    """256^2 Re 1000 to the 1e-5 residual: plain FP64, mixed, learned + mixed."""
    n = 256
    runs = {}
    for form in ("plain", "mixed", "learned"):
        s, its, conv, sec, log = _to_tolerance(_params(n, form),
                                               net if form == "learned" else None)
        runs[form] = (s, its, conv, sec, log)
        print(f"256^2 Re 1000 {form}: {its} outer iterations, {sec:.2f} s "
              f"({1e3 * sec / its:.3f} ms per iteration)", flush=True)
    yield runs
    for s, *_ in runs.values():
        s.close()


def test_learned_to_tolerance_256(runs_256):
    n = 256
    runs = runs_256
    for form, (s, its, conv, sec, log) in runs.items():
        assert conv, form
        m = log.max(axis=1)
        assert m[-1] <= 1e-5 and np.all(m[1:-1] > 1e-5), form  # the reference's stop rule
    assert abs(runs["learned"][1] - runs["mixed"][1]) <= 0.01 * runs["mixed"][1]
    p = runs["plain"][0].p
    H = p.H
    ref, lrn = runs["plain"][0], runs["learned"][0]
    errs = [float(np.abs(lrn.get(f) - ref.get(f)).max()) for f in ("u", "v")]
    pa = lrn.get("p")[H:H + n, H + 1:H + 1 + n]
    pb = ref.get("p")[H:H + n, H + 1:H + 1 + n]
    errs.append(float(np.abs((pa - pa.mean()) - (pb - pb.mean())).max()))
    kept, rej = sg.warm_stats(lrn)
    print(f"learned vs plain at tolerance: max|du| {errs[0]:.2e} max|dv| {errs[1]:.2e} "
          f"max|dp| {errs[2]:.2e}; guesses kept {kept}, discarded {rej}; time learned / plain "
          f"{runs['learned'][3] / runs['plain'][3]:.2f}", flush=True)
    assert max(errs) < 1e-6
    assert kept > 0.9 * runs["learned"][1]


def test_pressure_side_speedup_bound(runs_256):
    """What any p'-side method (the warm start included) can gain at most.

    The plain FP64 run solved with four V-cycles per p' solve instead of one (p' error per solve
    about 0.03 of the converged p' instead of 0.33-0.41) takes the same number of outer iterations
    to the tolerance (16,500 both; 20 V-cycles: 16,503): here the SIMPLE iteration count is set by
    the under-relaxation of u, v and p, not by how exactly p' is solved. So a p'-side method can
    at best make the p' solve free at the same iteration count, a speedup of at most
    1 / (1 - s), s = the p' solve's share of an outer iteration. s is taken from the hipEvent
    phase timers of the 1024^2 bench workload's eager iterations, which over-state it (the p' solve
    has the most and the shortest launches, so it carries the most launch gaps): the bound is
    generous, and still below 3."""
    plain_its = runs_256["plain"][1]
    p4 = _params(256, "plain")
    p4.mg_cycles = 4
    s, its4, conv, sec, _ = _to_tolerance(p4)
    s.close()
    print(f"256^2 plain, 4 V-cycles per p' solve: {its4} outer iterations ({sec:.1f} s), one "
          f"V-cycle: {plain_its}", flush=True)
    assert conv and abs(its4 - plain_its) <= 0.002 * plain_its
    q = _params(1024, "plain")
    b = sg.Solver(q)
    b.advance(3)
    b.set_phase_timing(True)
    b.advance(64)
    ph, timed = b.phase_times()
    b.close()
    share = ph[9] / ph.sum()  # solve_p (timer.hpp order)
    bound = (plain_its / its4) / (1.0 - share)
    print(f"1024^2 plain: p' solve {1e6 * ph[9] / timed:.0f} of {1e6 * ph.sum() / timed:.0f} us per "
          f"timed outer iteration (share {share:.2f}); largest speedup of a p'-side method "
          f"{bound:.2f}x", flush=True)
    assert timed > 0 and 0.0 < share < 1.0
    assert bound < 3.0
