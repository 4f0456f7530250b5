"""Invariants of the multigrid and CG p' solvers (CPU), restating the reference's
tests/unit/test_multigrid.cpp:75-268 and tests/unit/test_cg.cpp:80-187, plus the published
V-cycle factors of docs/RESULTS.md:141-142. A "realistic system" is the p' system assembled in the
5th gs_lex outer iteration from zero fields.
"""
import numpy as np
import pytest

import pyoracle as po
import simple_gpu as sg


def make_system(nx, ny, re, **kv):
    text = f"N_cells_x = {nx}\nN_cells_y = {ny}\nRe = {re}\n" + "".join(
        f"{k} = {v}\n" for k, v in kv.items())
    p, _ = sg.params_from_config(text)
    o = po.Oracle(p)
    for _ in range(5):
        o.outer_iteration()
    return p, o


def interior(p, a):
    H = p.H
    return a[H:p.NX + H, H + 1:p.NY + H + 1]


def true_rel_residual(p, o, x):
    ax = o.apply_A(x)
    b = interior(p, o.get("b"))
    bb = b - b.mean()
    r = bb - interior(p, ax)
    return np.sqrt((r * r).sum() / (bb * bb).sum())


@pytest.mark.parametrize("nx,ny,want", [
    (80, 80, [(80, 80), (40, 40), (20, 20), (10, 10), (5, 5)]),
    (32, 40, [(32, 40), (16, 20), (8, 10), (4, 5)]),
    (64, 64, [(64, 64), (32, 32), (16, 16), (8, 8), (4, 4)]),
    (512, 512, [(512, 512), (256, 256), (128, 128), (64, 64), (32, 32), (16, 16), (8, 8), (4, 4)]),
    (33, 64, [(33, 64)]),
])
def test_hierarchy_dimensions(nx, ny, want):
    p, _ = sg.params_from_config(f"N_cells_x = {nx}\nN_cells_y = {ny}\n")
    o = po.Oracle(p)
    assert [o.mg_dims(l) for l in range(o.mg_nlevels())] == want


def _stencil_apply(A, x):
    """A x on a level (arrays nx x ny: AE, AW, AN, AS, AP), zero outside."""
    AE, AW, AN, AS, AP = A
    xp = np.pad(x, 1)
    return AP * x - (AE * xp[1:-1, 2:] + AW * xp[1:-1, :-2] + AS * xp[2:, 1:-1]
                     + AN * xp[:-2, 1:-1])


def test_galerkin_coarse_operator():
    p, o = make_system(32, 40, 400)
    o.mg_build()
    rng = np.random.default_rng(12345)
    for l in range(1, o.mg_nlevels()):
        F = [o.mg_get(l - 1, k) for k in range(5)]
        Cc = [o.mg_get(l, k) for k in range(5)]
        xc = rng.uniform(-1, 1, size=Cc[0].shape)
        ac = _stencil_apply(Cc, xc)
        xf = np.kron(xc, np.ones((2, 2)))
        af = _stencil_apply(F, xf)
        rac = af[0::2, 0::2] + af[0::2, 1::2] + af[1::2, 0::2] + af[1::2, 1::2]
        amax = np.abs(ac).max()
        assert amax > 0 and np.abs(ac - rac).max() <= 1e-12 * amax


def test_coarse_rows_sum_zero_symmetric_walls_zero():
    p, o = make_system(32, 40, 400)
    o.mg_build()
    for l in range(1, o.mg_nlevels()):
        AE, AW, AN, AS, AP = [o.mg_get(l, k) for k in range(5)]
        assert np.all(AP > 0)
        assert np.all(np.abs(AP - (AE + AW + AN + AS)) <= 1e-14 * AP)
        assert np.all(np.abs(AE[:, :-1] - AW[:, 1:]) <= 1e-14 * np.abs(AE[:, :-1]))
        assert np.all(np.abs(AS[:-1, :] - AN[1:, :]) <= 1e-14 * np.abs(AS[:-1, :]))
        assert np.all(AE[:, -1] == 0) and np.all(AS[-1, :] == 0)
        assert np.all(AW[:, 0] == 0) and np.all(AN[0, :] == 0)


def test_vcycle_convergence_and_match_with_cg():
    p, o = make_system(64, 64, 1000)
    o.mg_build()
    o.set("pc", np.zeros(o.shape))
    rel = [1.0]
    k1e8 = -1
    for k in range(1, 61):
        o.mg_vcycle()
        rel.append(true_rel_residual(p, o, o.get("pc")))
        if k1e8 < 0 and rel[-1] <= 1e-8:
            k1e8 = k
    factor10 = (rel[10] / rel[0]) ** 0.1
    assert factor10 < 0.5
    assert 0 < k1e8 <= 40
    x_mg = interior(p, o.get("pc"))
    x_mg = x_mg - x_mg.mean()
    q = p.copy()
    q.pressure_tolerance = 1e-12
    o.set_params(q)
    o.solve_p_kind(sg.P_CG)
    assert o.cg_last_rel() <= 1e-12
    x_cg = interior(p, o.get("pc"))
    assert np.abs(x_cg).max() > 1e-6
    assert np.abs(x_mg - x_cg).max() <= 1e-8


@pytest.mark.parametrize("n,f18,f1", [(64, 0.255, 0.796), (128, 0.301, 0.816)])
def test_published_vcycle_factors(n, f18, f1):
    """docs/RESULTS.md:141-142: average reduction per V-cycle over 10 cycles from p' = 0."""
    got = {}
    for omega in (1.0, 1.8):
        p, o = make_system(n, n, 1000, mg_omega=omega)
        o.mg_build()
        o.set("pc", np.zeros(o.shape))
        for _ in range(10):
            o.mg_vcycle()
        got[omega] = true_rel_residual(p, o, o.get("pc")) ** 0.1
    assert abs(got[1.8] - f18) < 5e-4 and abs(got[1.0] - f1) < 5e-4, got
    assert got[1.8] < 0.5 and got[1.8] < got[1.0]


def test_nonzero_mean_rhs():
    p, o = make_system(64, 64, 1000, mg_cycles=30)
    o.solve_p_kind(sg.P_MG)
    x0 = interior(p, o.get("pc"))
    b = o.get("b")
    scale = np.abs(interior(p, b)).max()
    interior(p, b)[...] += 0.5 * scale
    o.set("b", b)
    o.solve_p_kind(sg.P_MG)
    x = interior(p, o.get("pc"))
    assert np.all(np.isfinite(x))
    assert abs(x.mean()) <= 1e-12 * np.abs(x).max()
    assert np.abs(x - x0).max() <= 1e-4 * np.abs(x0).max()


# ---- CG (tests/unit/test_cg.cpp) ----

def test_cg_reaches_tolerance():
    p, o = make_system(32, 32, 1000)
    o.solve_p_kind(sg.P_CG)
    k = o.cg_last_iterations()
    assert 0 < k < p.pressure_maxit
    assert o.cg_last_rel() <= p.pressure_tolerance
    assert true_rel_residual(p, o, o.get("pc")) <= 2 * p.pressure_tolerance
    x = interior(p, o.get("pc"))
    assert abs(x.mean()) <= 1e-13 * np.abs(x).max()


def test_cg_matches_gauss_seidel():
    p, o = make_system(32, 32, 1000)
    q = p.copy()
    q.pressure_tolerance = 1e-12
    o.set_params(q)
    b = o.get("b")
    o.solve_p_kind(sg.P_CG)
    x_cg = interior(p, o.get("pc"))
    bi = interior(p, b)
    bi -= bi.mean()
    o.set("b", b)
    q2 = q.copy()
    q2.maxit_inner_p = 20000
    o.set_params(q2)
    o.solve_p_kind(sg.P_GS_LEX)
    x_gs = interior(p, o.get("pc"))
    x_gs = x_gs - x_gs.mean()
    assert np.abs(x_cg).max() > 1e-6
    assert np.abs(x_cg - x_gs).max() <= 1e-8


def test_cg_a_norm_error_non_increasing():
    p, o = make_system(32, 32, 1000)
    q = p.copy()
    q.pressure_tolerance = 1e-13
    o.set_params(q)
    o.solve_p_kind(sg.P_CG)
    xs = o.get("pc")
    q.pressure_tolerance = 1e-6
    o.set_params(q)
    o.solve_p_kind(sg.P_CG)
    K = o.cg_last_iterations()
    prev = np.inf
    for k in range(1, K + 1, max(1, K // 25)):
        q.pressure_maxit = k
        o.set_params(q)
        o.solve_p_kind(sg.P_CG)
        e = np.zeros(o.shape)
        interior(p, e)[...] = interior(p, o.get("pc")) - interior(p, xs)
        en = (interior(p, e) * interior(p, o.apply_A(e))).sum()
        assert en <= prev * (1 + 1e-10)
        prev = en


def test_cg_zero_rhs_gives_zero():
    p, o = make_system(32, 32, 1000)
    o.set("b", np.zeros(o.shape))
    o.set("pc", np.ones(o.shape))
    o.solve_p_kind(sg.P_CG)
    assert o.cg_last_iterations() == 0
    assert np.abs(interior(p, o.get("pc"))).max() == 0.0


@pytest.mark.parametrize("dt,g,rng_lo,rng_hi", [(np.float64, 1e-12, 1e-280, 1e280),
                                                (np.float32, 1e-5, 1e-30, 1e30)])
def test_coarse_stop_test_bounds(dt, g, rng_lo, rng_hi):
    """The coarse solve's stop test sqrt(rr) < c decides by comparing rr with t (1 -/+ g),
    t = c * c rounded, and takes the root only between those bounds (or when t leaves the normal
    range). Wherever the comparison decides, it must agree with the correctly rounded root: rr
    right at and around t, the bounds and their neighbours, and random values, over c spread
    across the exponent range."""
    rs = np.random.default_rng(3)
    one = dt(1)
    emax = 300 if dt is np.float64 else 36
    cs = (10.0 ** rs.uniform(-emax / 2, emax / 2, 400)).astype(dt)
    cs = np.concatenate([cs, np.array([1, 2, 0.5, 1e-10, 3.0], dt)])
    decided = 0
    for c in cs:
        t = dt(c * c)
        if not (rng_lo < t < rng_hi):
            continue
        lo, hi = dt(t * (one - dt(g))), dt(t * (one + dt(g)))
        probes = [t, lo, hi]
        for b in (t, lo, hi):
            x = b
            y = b
            for _ in range(40):
                x, y = np.nextafter(x, dt(np.inf)), np.nextafter(y, dt(0))
                probes += [x, y]
        probes += list((t * rs.uniform(0.5, 2.0, 50)).astype(dt))
        probes += list((t * (1 + rs.uniform(-4, 4, 50) * g)).astype(dt))
        for rr in np.array(probes, dt):
            root = bool(np.sqrt(rr) < c)
            if rr < lo:
                assert root, (c, rr)
                decided += 1
            elif rr > hi:
                assert not root, (c, rr)
                decided += 1
    assert decided > 10000
