"""Pins of the CPU oracle (CPU only).

(i)  the reference's golden snapshots (tests/snapshots/{short_16,converged_32x40}, copied under
     tests/golden/snapshots): u, v, p bit-identical, same iteration count and last alog row
     (the reference's tests/regression/test_snapshots.cpp:11-40);
(ii) published outer-iteration counts of docs/RESULTS.md (a bit-identical port must hit them
     exactly; a dropped term, a wrong sign or a different summation order would not);
(iii) the reference itself compiled from its sources into oracle/_ref (when present in this
     container): every field after every phase, for every solver combination;
(iv) closed forms: res_v == 0 exactly at iteration 1 (tests/unit/test_boundary.cpp:67) and the
     first res_u of short_16 from zero fields with lid ghost 2.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, bits_equal, first_diff, golden_cfg, load_csv, with_keys

import pyoracle as po
import simple_gpu as sg


def _oracle(cfg):
    p, _ = sg.params_from_config(cfg)
    return p, po.Oracle(p)


@pytest.mark.parametrize("name", ["short_16", "converged_32x40"])
def test_snapshot_bit_exact(name):
    p, o = _oracle(golden_cfg(name))
    n, log, conv = o.run()
    d = os.path.join(GOLDEN, "snapshots", name)
    for f in ["u", "v", "p"]:
        want = load_csv(os.path.join(d, f + ".csv"))
        assert bits_equal(o.get(f), want), f"{f}: {first_diff(o.get(f), want)}"
    alog = np.loadtxt(os.path.join(d, "alog.csv"), delimiter=",", skiprows=1, ndmin=2)
    assert len(alog) == n
    assert bits_equal(alog[-1], log[-1])
    assert bits_equal(alog, log)


def test_first_iteration_closed_forms():
    p, o = _oracle(golden_cfg("short_16"))
    r = o.outer_iteration()
    # From zero fields only the lid ghost (2 - 0) drives u; v stays exactly 0 (test_boundary.cpp:67).
    assert r[1] == 0.0
    assert r[0] == 0.3233333333333333
    o2 = po.Oracle(p)
    o2.assemble_u()
    o2.bc_u()
    u = o2.get("u")
    H = p.H
    assert np.all(u[H - 1, H + 1:p.NY + H + 1] == 2.0)


PUBLISHED = [
    # (label, config overrides, outer iterations) — docs/RESULTS.md:69-77, :149, :151
    ("32 Re400 gs_lex/gs_lex", dict(momentum_solver="gs_lex", pressure_solver="gs_lex"), 1404),
    ("32 Re400 jacobi/jacobi", dict(momentum_solver="jacobi", pressure_solver="jacobi"), 2369),
    ("32 Re400 mc/rb", dict(momentum_solver="gs_multicolour", pressure_solver="gs_redblack"), 1470),
    ("32 Re400 mc/gs_lex", dict(momentum_solver="gs_multicolour", pressure_solver="gs_lex"), 1456),
    ("32 Re400 gs_lex/rb", dict(momentum_solver="gs_lex", pressure_solver="gs_redblack"), 1423),
    ("32 Re400 jacobi/gs_lex", dict(momentum_solver="jacobi", pressure_solver="gs_lex"), 1598),
    ("32 Re400 gs_lex/jacobi", dict(momentum_solver="gs_lex", pressure_solver="jacobi"), 2186),
    ("32 Re400 gs_lex/mg", dict(momentum_solver="gs_lex", pressure_solver="mg"), 894),
    ("32 Re400 gs_lex/mg x2", dict(momentum_solver="gs_lex", pressure_solver="mg", mg_cycles=2), 883),
    ("32 Re400 gs_lex/mg w1", dict(momentum_solver="gs_lex", pressure_solver="mg", mg_omega=1), 1049),
    ("32 Re400 gs_lex/mg w1 x2", dict(momentum_solver="gs_lex", pressure_solver="mg", mg_omega=1,
                                      mg_cycles=2), 954),
]


@pytest.mark.parametrize("label,kv,want", PUBLISHED, ids=[c[0] for c in PUBLISHED])
def test_published_iteration_counts_32(label, kv, want):
    cfg = "N_cells_x = 32\nN_cells_y = 32\nRe = 400\nerror_tolerance = 1e-9\n"
    p, o = _oracle(with_keys(cfg, **kv))
    n, _, conv = o.run()
    assert conv and n == want


VALIDATION = ("N_cells_x = {n}\nN_cells_y = {n}\nRe = {re}\nmaxit = 60000\nerror_tolerance = 1e-5\n"
              "convection_scheme = QUICK\ndiffusion_order = 4\n")


@pytest.mark.parametrize("n,re,mom,pres,want", [
    (80, 1000, "gs_multicolour", "gs_redblack", 3718),   # docs/RESULTS.md:79
    (80, 400, "gs_multicolour", "gs_redblack", 2337),    # :79
    (128, 1000, "gs_multicolour", "mg", 4703),           # :169
])
def test_published_iteration_counts_validation(n, re, mom, pres, want):
    po.set_threads(min(8, os.cpu_count() or 1))
    p, o = _oracle(with_keys(VALIDATION.format(n=n, re=re), momentum_solver=mom,
                             pressure_solver=pres))
    k, _, conv = o.run()
    assert conv and k == want


def test_thread_invariance():
    """Bit-identical fields for 1 and 8 threads (tests/regression/test_threads.cpp:16-51)."""
    cfg = with_keys(golden_cfg("converged_32x40"), momentum_solver="gs_multicolour",
                    pressure_solver="mg", maxit=60, error_tolerance=0)
    out = []
    for t in (1, 8):
        po.set_threads(t)
        p, o = _oracle(cfg)
        o.run()
        out.append([o.get(f) for f in ("u", "v", "p")])
    po.set_threads(os.cpu_count() or 1)
    for a, b in zip(*out):
        assert bits_equal(a, b)


# ---- against the reference's own code (oracle/_ref), when it is built here ----

needs_ref = pytest.mark.skipif(not po.ref_available(),
                               reason="oracle/_ref not built (reference sources absent)")

COMBOS = [(m, q) for m in ("gs_lex", "jacobi", "gs_multicolour")
          for q in ("gs_lex", "jacobi", "gs_redblack", "cg", "mg")]


@needs_ref
@pytest.mark.parametrize("scheme,order", [("QUICK", 4), ("UDS", 2)])
@pytest.mark.parametrize("mom,pres", COMBOS)
def test_against_reference_build(mom, pres, scheme, order):
    cfg = with_keys(golden_cfg("converged_32x40"), momentum_solver=mom, pressure_solver=pres,
                    convection_scheme=scheme, diffusion_order=order, maxit=25, error_tolerance=0)
    p, o = _oracle(cfg)
    r = po.Reference(cfg)
    for it in range(25):
        a, b = o.outer_iteration(), r.outer_iteration()
        assert bits_equal(a, b), (it, a, b)
    for f in po.FIELDS:
        assert bits_equal(o.get(f), r.get(f)), f"{f}: {first_diff(o.get(f), r.get(f))}"


@needs_ref
def test_against_reference_per_phase_random_state():
    """Per phase from a random state (LCG-filled u, v, p): catches order-of-operation slips that
    the smooth states of a run might hide."""
    cfg = with_keys(golden_cfg("converged_32x40"), momentum_solver="gs_multicolour",
                    pressure_solver="mg", maxit=1, error_tolerance=0)
    p, o = _oracle(cfg)
    r = po.Reference(cfg)
    rng = np.random.default_rng(12345)
    for f in ("u", "v", "p"):
        a = rng.uniform(-0.5, 0.5, size=o.shape)
        o.set(f, a)
        r.set(f, a)
    steps = [("assemble_u", ()), ("bc_u", ()), ("residual_u", ()), ("sweep_u", (4,)),
             ("assemble_v", ()), ("bc_v", ()), ("residual_v", ()), ("sweep_v", (4,)),
             ("assemble_p", ()), ("solve_p", ()), ("correct", ())]
    for name, args in steps:
        ra = getattr(o, name)(*args)
        rb = getattr(r, name)(*args)
        if ra is not None:
            assert ra == rb, (name, ra, rb)
        for f in po.FIELDS:
            assert bits_equal(o.get(f), r.get(f)), f"{name} {f}: {first_diff(o.get(f), r.get(f))}"


@needs_ref
def test_params_derivation_matches_reference():
    for cfg in ["N_cells_x = 32\nN_cells_y = 40\nRe = 400\n", "N_cells_x = 7\nN_cells_y = 3\n",
                "N_cells_x = 1024\nN_cells_y = 1024\nRe = 3200\nN_halo_x = 3\nN_halo_y = 3\n"]:
        p, _ = sg.params_from_config(cfg)
        r = po.Reference(cfg)
        assert r.derived == (p.nrows, p.ncols, p.dx, p.dy, p.nu)


@pytest.mark.skipif(not po.ref_available(), reason="the reference build oracle/_ref is absent")
def test_bench_cpu_baseline_drives_the_reference():
    """bench.py's cpu_baseline times the reference's own CPU code: the workload's config text with
    this backend's added keys dropped parses in the reference's parser (which rejects unknown
    keys), and its outer iterations give the oracle's residuals bit for bit."""
    import sys
    from conftest import ROOT
    sys.path.insert(0, ROOT)
    import bench
    o = bench.CpuSolver(sg, bench.config_text(32, "learned"), 2)  # every key bench writes
    assert o.kind == "reference"
    _, ref = _oracle(bench.config_text(32, "plain"))
    for _ in range(3):
        a, b = o.o.outer_iteration(), ref.outer_iteration()
        assert np.array_equal(a, b), (a, b)
    assert o.step() > 0.0
