"""The workloads bench.py times, checked against the CPU oracle at their full sizes.

 - 1024 x 1024, Re 1000, QUICK / 4th-order, multicolour momentum, multigrid V(2,2) p' (the bench
   workload): the FP64 form through graph replays (iterations 1-3 in one-iteration graphs, then
   one eight-iteration graph and three one-iteration graphs), with the bench's fixed launch
   shapes and with autotuned ones, bit for bit against the oracle: the residual log and every
   array (u, v, p, p', b, all 18 momentum coefficients and the 5 p' operator arrays);
 - the mixed forms at 1024^2 against the FP64 oracle, the error of u, v and p printed and bounded;
 - the 64 uneven cavities of the batched workload (workloads/cavities_seed0.txt), one persistent
   launch per phase, every cavity against its own oracle run (FP64 bit for bit; mixed within
   tolerance), and per-cavity stop rules against the oracle's iteration counts;
 - 8192 x 8192, Re 3200 (the large workload): two FP64 outer iterations bit for bit against the
   oracle, and the mixed form's error.
"""
import numpy as np
import pytest

from conftest import bits_equal, first_diff

pytestmark = pytest.mark.gpu

import pyoracle as po  # noqa: E402
import simple_gpu as sg  # noqa: E402

# The bench workload's config text (bench.py config_text) and its fixed launch shapes (--knobs).
BENCH_KNOBS = (11, 1024, 0, 4096, 0, 5)


def _bench_cfg(n, form="plain", re=1000, maxit=10000000, tol="1e-5", order=4):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = {re}\nconvection_scheme = QUICK\n"
            f"diffusion_order = {order}\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_pre = 2\nmg_post = 2\nerror_tolerance = {tol}\nmaxit = {maxit}\nbackend = gpu\n"
            f"write_fields = False\ncheck_every = 500\n")
    if form in ("mixed", "mixed_mom"):
        text += "mg_precision = mixed\n"
    if form == "mixed_mom":
        text += "momentum_precision = mixed\n"
    return sg.params_from_config(text)[0]


def _same(name, g, c):
    assert bits_equal(g, c), f"{name}: {first_diff(g, c)}"


def _interior(p, a):
    return a[p.H:p.NX + p.H, p.H + 1:p.NY + p.H + 1]


def _errors(p, g, cpu):
    du = float(np.abs(g.get("u") - cpu.get("u")).max())
    dv = float(np.abs(g.get("v") - cpu.get("v")).max())
    pg, pc = _interior(p, g.get("p")), _interior(p, cpu.get("p"))
    dp = float(np.abs((pg - pg.mean()) - (pc - pc.mean())).max())
    return du, dv, dp


@pytest.fixture(scope="module")
def oracle_1024():
    """The oracle after 3 and after 14 FP64 outer iterations of the bench workload."""
    po.set_threads(max(1, min(16, __import__("os").cpu_count() or 1)))
    p = _bench_cfg(1024)
    cpu = po.Oracle(p)
    logs = [cpu.outer_iteration() for _ in range(14)]
    state = {f: cpu.get(f) for f in sg.FIELDS}
    return p, np.array(logs), state


@pytest.mark.parametrize("tuning", ["fixed", "autotuned"])
def test_bench_workload_1024_bit_identical(oracle_1024, tuning):
    p, logc, ref = oracle_1024
    g = sg.Solver(p)
    if tuning == "fixed":
        sg.set_tunings(g, BENCH_KNOBS)
    print(f"1024^2 bench workload, {tuning} launch shapes", flush=True)
    d1, l1, _ = g.advance(3)
    if tuning == "autotuned":  # as bench.py: autotune on the running state, results unchanged
        rep = g.autotune(reps=2)
        print(rep.strip().splitlines()[-1], "knobs", sg.tuning_of(g), flush=True)
    d2, l2, _ = g.advance(11)  # one 8-iteration graph and three 1-iteration graphs
    assert d1 == 3 and d2 == 11 and g.iterations() == 14
    logg = np.concatenate([l1, l2])
    assert bits_equal(logg, logc), first_diff(logg, logc)
    for f in sg.FIELDS:
        _same(f"1024^2 {tuning} {f}", g.get(f), ref[f])
    g.close()


# Bounds about 10x the measured errors after 60 iterations: mixed multigrid 1.2e-9 / 9.0e-10 /
# 2.1e-9 for u / v / p; with the FP32 momentum corrections too, 5.0e-9 / 3.0e-9 / 6.1e-9.
@pytest.mark.parametrize("form,bound", [("mixed", 2e-8), ("mixed_mom", 6e-8)])
def test_bench_workload_1024_mixed_error(form, bound):
    """The mixed forms against the FP64 oracle after 60 outer iterations from zero fields: the
    error of u, v and p (p up to its mean, which the method leaves free) printed and bounded."""
    n = 60
    p64 = _bench_cfg(1024)
    cpu = po.Oracle(p64)
    cpu.run_fixed(n)
    pm = _bench_cfg(1024, form)
    g = sg.Solver(pm)
    done, _, _ = g.advance(n)
    assert done == n
    du, dv, dp = _errors(pm, g, cpu)
    print(f"\n1024^2 {form} after {n} outer iterations: max|du| {du:.2e} max|dv| {dv:.2e} "
          f"max|dp| {dp:.2e}", flush=True)
    assert max(du, dv, dp) < bound, (du, dv, dp)
    g.close()


def _batch_params(cases, precision, tol=0.0, maxit=100000):
    return [sg.params_from_config(sg.batch_config(n, re, precision=precision, tol=tol,
                                                  maxit=maxit, check_every=50))[0]
            for n, re in cases]


def test_batch_workload_64_cavities_vs_oracle():
    """All 64 cavities of the batched workload, 7 FP64 outer iterations in one batch, every
    cavity's u, v, p, p' and residual bit for bit against its own oracle run."""
    cases = sg.cavities()
    assert len(cases) == 64
    iters = 7
    b = sg.Batch(_batch_params(cases, "fp64"))
    print(f"batch of {len(cases)} cavities, {b.launches()} launches per iteration", flush=True)
    its, st, last = b.advance(iters)
    assert its == [iters] * len(cases) and st == [0] * len(cases)
    for k, (n, re) in enumerate(cases):
        cpu = po.Oracle(b.plist[k])
        logc = [cpu.outer_iteration() for _ in range(iters)]
        c = b.cavity(k)
        assert bits_equal(last[k], logc[-1]), (k, n, re, last[k], logc[-1])
        for f in ("u", "v", "p", "pc"):
            _same(f"cavity {k} ({n}, Re {re}) {f}", c.get(f), cpu.get(f))
    b.close()


def test_batch_workload_64_cavities_mixed_vs_oracle():
    """The same 64 cavities in the mixed form (FP32 V-cycle inside FP64 defect correction), 40
    outer iterations, each within 3e-7 of its FP64 oracle run (about 10x the measured worst)."""
    cases = sg.cavities()
    iters = 40
    b = sg.Batch(_batch_params(cases, "mixed"))
    its, _, _ = b.advance(iters)
    assert its == [iters] * len(cases)
    worst = 0.0
    for k, (n, re) in enumerate(cases):
        cpu = po.Oracle(_batch_params([(n, re)], "fp64")[0])
        cpu.run_fixed(iters)
        e = _errors(b.plist[k], b.cavity(k), cpu)
        worst = max(worst, max(e))
        assert max(e) < 3e-7, (k, n, re, e)  # measured worst: 3.2e-8
    print(f"\n64 cavities mixed vs oracle after {iters} iterations: worst max error {worst:.2e}",
          flush=True)
    b.close()


def test_batch_stop_rules_match_oracle():
    """Uneven cavities to tolerance in one batch: each stops at the oracle's iteration count with
    the oracle's fields."""
    cases = [(16, 150.0), (24, 600.0), (32, 300.0), (20, 1000.0), (48, 2500.0), (40, 120.0)]
    plist = _batch_params(cases, "fp64", tol=1e-4, maxit=30000)
    b = sg.Batch(plist)
    its, st, _ = b.advance(30000)
    print("batch iterations", its, "status", st, flush=True)
    assert st == [1] * len(cases) and len(set(its)) > 1
    for k, p in enumerate(plist):
        cpu = po.Oracle(p)
        nc, logc, conv = cpu.run()
        assert conv and its[k] == nc, (k, cases[k], its[k], nc)
        c = b.cavity(k)
        for f in ("u", "v", "p"):
            _same(f"cavity {k} {f}", c.get(f), cpu.get(f))
    b.close()


def test_large_8192_re3200_bit_identical():
    """The large workload (8192^2, Re 3200, QUICK / 4th order, multigrid): two FP64 outer
    iterations (~16 GB on the device, as much again on the host) bit for bit against the oracle,
    then the mixed form's error after the same two iterations."""
    po.set_threads(max(1, min(16, __import__("os").cpu_count() or 1)))
    p = _bench_cfg(8192, re=3200)
    cpu = po.Oracle(p)
    logc = np.array([cpu.outer_iteration() for _ in range(2)])
    print("8192^2: oracle done", flush=True)
    g = sg.Solver(p)
    done, logg, _ = g.advance(2)
    assert done == 2
    assert bits_equal(logg, logc), first_diff(logg, logc)
    for f in ("u", "v", "p", "pc", "b"):
        _same("8192^2 " + f, g.get(f), cpu.get(f))
    g.close()
    pm = _bench_cfg(8192, "mixed", re=3200)
    g = sg.Solver(pm)
    g.advance(2)
    du, dv, dp = _errors(pm, g, cpu)
    pc = _interior(pm, cpu.get("p"))
    mu, mv, mp = (float(np.abs(cpu.get("u")).max()), float(np.abs(cpu.get("v")).max()),
                  float(np.abs(pc - pc.mean()).max()))
    print(f"\n8192^2 mixed after 2 iterations: max|du| {du:.2e} max|dv| {dv:.2e} "
          f"max|dp| {dp:.2e}; relative to max|u| {mu:.3g}, max|v| {mv:.3g}, max|p| {mp:.3g}: "
          f"{du / mu:.2e} {dv / mv:.2e} {dp / mp:.2e}", flush=True)
    # Two iterations from zero fields: p' is at its largest, and one FP32 V-cycle carries an
    # error of about 3e-7 of it (the mixed solve test), so the bound is relative.
    assert du / mu < 2e-6 and dv / mv < 2e-6 and dp / mp < 2e-6, (du, dv, dp)
    g.close()
