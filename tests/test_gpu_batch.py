"""Batched cavities: one persistent launch per phase over cavities of uneven size and Reynolds
number must give every cavity exactly what it gets alone.

The solo runs are the single-cavity sg_advance path, itself bit-identical to the oracle (the
per-phase and whole-run parity tests). Here every cavity of a batch is compared bit for bit with
its own solo run (fields and iteration counts), for the fp64 and mixed multigrid forms and the
red-black pressure solver; the stop rules must fire per cavity; the number of launches of a
batched iteration must not grow with the number of cavities; and the mixed-precision batch must
stay within 1e-6 of the fp64 batch on u, v and p at the same tolerance.
"""
import numpy as np
import pytest

from conftest import bits_equal, first_diff

pytestmark = pytest.mark.gpu

import simple_gpu as sg  # noqa: E402

# Small uneven cavities: multigrid depths 2..5, Re spread over the workload's range.
SMALL = [(16, 150.0), (24, 2900.0), (32, 700.0), (40, 1200.0), (64, 400.0), (20, 3100.0)]


def _params(cases, precision="fp64", pressure="mg", tol=0.0, maxit=100000):
    out = []
    for n, re in cases:
        text = sg.batch_config(n, re, precision=precision, tol=tol, maxit=maxit, check_every=50)
        text = text.replace("pressure_solver = mg", f"pressure_solver = {pressure}")
        p, _ = sg.params_from_config(text)
        out.append(p)
    return out


def _solo(p, iters):
    s = sg.Solver(p)
    s.advance(iters)
    r = {f: s.get(f) for f in ("u", "v", "p")}
    r["iters"] = s.iterations()
    s.close()
    return r


@pytest.mark.parametrize("precision,pressure", [("fp64", "mg"), ("mixed", "mg"),
                                                ("fp64", "gs_redblack")])
def test_batch_matches_solo(precision, pressure):
    plist = _params(SMALL, precision, pressure)
    print(f"batch {precision}/{pressure}: {len(plist)} cavities, 37 iterations", flush=True)
    b = sg.Batch(plist)
    # Coarse levels above the one-workgroup tail for the larger cavities only, so the batch merges
    # multigrid sequences of different depths (launch shapes do not change results).
    b.set_tuning(sg.TUNE_MID_CELLS, 0)
    b.set_tuning(sg.TUNE_LDS_CELLS, 256)
    its, st, _ = b.advance(37)
    assert its == [37] * len(plist)
    for k, p in enumerate(plist):
        ref = _solo(p, 37)
        c = b.cavity(k)
        for f in ("u", "v", "p"):
            g = c.get(f)
            assert bits_equal(g, ref[f]), f"cavity {k} {SMALL[k]} {f}: {first_diff(g, ref[f])}"
    b.close()


def test_batch_stop_rules_per_cavity():
    cases = [(16, 150.0), (24, 600.0), (32, 300.0), (20, 1000.0)]
    plist = _params(cases, "fp64", "mg", tol=1e-4)
    b = sg.Batch(plist)
    its, st, last = b.advance(20000)
    print("batch iterations", its, "status", st, flush=True)
    assert st == [1] * len(cases)
    assert len(set(its)) > 1  # the cavities stop at different iterations
    for k, p in enumerate(plist):
        ref = _solo(p, 20000)
        assert its[k] == ref["iters"], (k, its[k], ref["iters"])
        c = b.cavity(k)
        for f in ("u", "v", "p"):
            assert bits_equal(c.get(f), ref[f]), (k, f)
        assert max(last[k]) <= 1e-4
    b.close()


def test_batch_launches_independent_of_count():
    a = sg.Batch(_params([(16, 100.0), (64, 500.0)]))
    b = sg.Batch(_params([(16, 100.0), (64, 500.0), (16, 900.0), (64, 2000.0), (16, 3000.0)]))
    la, lb = a.launches(), b.launches()
    print("launches per batched iteration:", la, lb, flush=True)
    assert la == lb
    a.close()
    b.close()


def test_batch_autotune_keeps_results():
    plist = _params(SMALL[:4], "mixed", "mg")
    b = sg.Batch(plist)
    b.advance(5)
    rep = b.autotune(reps=3)
    print(rep, flush=True)
    assert "best momentum tile" in rep and rep.strip().splitlines()[-1].startswith("best:")
    b.advance(10)
    for k, p in enumerate(plist):
        ref = _solo(p, 15)
        c = b.cavity(k)
        for f in ("u", "v", "p"):
            assert bits_equal(c.get(f), ref[f]), (k, f)
    b.close()


def test_batch_tuning_keys_round_trip():
    """Every launch-shape knob bench.py passes (sg.TUNE_KEYS, the FP32 correction tile included)
    reads and writes on a batch as on a solver."""
    b = sg.Batch(_params(SMALL[:2], "mixed", "mg"))
    t = sg.tuning_of(b)
    assert len(t) == len(sg.TUNE_KEYS)
    sg.set_tunings(b, t)
    assert sg.tuning_of(b) == t
    b.set_tuning(sg.TUNE_DC_TILE, 2)
    assert b.get_tuning(sg.TUNE_DC_TILE) == 2
    b.close()


def test_batch_mixed_close_to_fp64_at_tolerance():
    # The workload's first cavities at a quarter of their size (same Re draws), run to 1e-5.
    cases = [(n // 4, re) for n, re in sg.cavities()[:6]]
    mixed = sg.Batch(_params(cases, "mixed", "mg", tol=1e-5))
    fp64 = sg.Batch(_params(cases, "fp64", "mg", tol=1e-5))
    im, sm, _ = mixed.advance(200000)
    i6, s6, _ = fp64.advance(200000)
    assert sm == [1] * len(cases) and s6 == [1] * len(cases)
    for k in range(len(cases)):
        cm, c6 = mixed.cavity(k), fp64.cavity(k)
        errs = [float(np.max(np.abs(cm.get(f) - c6.get(f)))) for f in ("u", "v", "p")]
        print(f"cavity {cases[k]}: iterations mixed {im[k]} fp64 {i6[k]}; max|du| {errs[0]:.2e} "
              f"max|dv| {errs[1]:.2e} max|dp| {errs[2]:.2e}", flush=True)
        assert max(errs) < 1e-6, (k, errs)
    mixed.close()
    fp64.close()
