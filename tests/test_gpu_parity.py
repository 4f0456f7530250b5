"""GPU parity: the HIP backend against the CPU oracle, through the C ABI.

Same shape as the reference's GPU test (tests/hip/test_hip.cpp:70-183): non-trivial starting
fields from 3 CPU outer iterations on the 32x40 snapshot case, then each phase on the CPU and the
GPU in turn, every array compared bit for bit; residuals compared to 1e-12 relative (they are in
fact summed in the CPU's order, so they come out identical); whole runs compared bit for bit;
and run-to-run determinism.
"""
import numpy as np
import pytest

from conftest import bits_equal, first_diff, golden_cfg, with_keys

pytestmark = pytest.mark.gpu

import pyoracle as po  # noqa: E402
import simple_gpu as sg  # noqa: E402


def _cfg(scheme, order, mom, pres, maxit, **extra):
    return with_keys(golden_cfg("converged_32x40"), convection_scheme=scheme,
                     diffusion_order=order, momentum_solver=mom, pressure_solver=pres,
                     maxit=maxit, error_tolerance=0, backend="gpu", **extra)


def _same(name, g, c):
    assert bits_equal(g, c), f"{name}: {first_diff(g, c)}"


def _rel_close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(abs(b), 1e-300)


COEFS = ["AE", "AW", "AN", "AS", "AEE", "AWW", "ANN", "ASS", "AP"]


def _per_phase(scheme, order, mom, pres, **extra):
    p, _ = sg.params_from_config(_cfg(scheme, order, mom, pres, 3, **extra))
    cpu = po.Oracle(p)
    for _ in range(3):
        cpu.outer_iteration()
    gpu = sg.Solver(p)
    for name in sg.FIELDS:
        gpu.set(name, cpu.get(name))
    print(f"per-phase {scheme}/{order} {mom}/{pres}", flush=True)

    cpu.assemble_u(); gpu.assemble_u()
    for c in COEFS:
        _same("assemble_u " + c, gpu.get("cu_" + c), cpu.get("cu_" + c))
    cpu.bc_u(); gpu.bc_u()
    _same("bc_u", gpu.get("u"), cpu.get("u"))
    rc, rg = cpu.residual_u(), gpu.residual_u()
    assert _rel_close(rg, rc), (rg, rc)
    cpu.sweep_u(p.maxit_inner_u); gpu.sweep_u(p.maxit_inner_u)
    _same("sweep_u", gpu.get("u"), cpu.get("u"))

    cpu.assemble_v(); gpu.assemble_v()
    for c in COEFS:
        _same("assemble_v " + c, gpu.get("cv_" + c), cpu.get("cv_" + c))
    cpu.bc_v(); gpu.bc_v()
    _same("bc_v", gpu.get("v"), cpu.get("v"))
    rc, rg = cpu.residual_v(), gpu.residual_v()
    assert _rel_close(rg, rc), (rg, rc)
    cpu.sweep_v(p.maxit_inner_v); gpu.sweep_v(p.maxit_inner_v)
    _same("sweep_v", gpu.get("v"), cpu.get("v"))

    rc, rg = cpu.assemble_p(), gpu.assemble_p()
    assert _rel_close(rg, rc), (rg, rc)
    for n in ["AEP", "AWP", "ANP", "ASP", "APP", "b"]:
        _same("assemble_p " + n, gpu.get(n), cpu.get(n))
    cpu.solve_p(); gpu.solve_p()
    _same("solve_p pc", gpu.get("pc"), cpu.get("pc"))
    _same("solve_p b", gpu.get("b"), cpu.get("b"))  # mg removes the mean of b in place
    cpu.correct(); gpu.correct()
    for n in ["u", "v", "p"]:
        _same("correct " + n, gpu.get(n), cpu.get(n))
    gpu.close()


@pytest.mark.parametrize("scheme,order", [("QUICK", 4), ("UDS", 2), ("QUICK", 2), ("UDS", 4)])
def test_per_phase_multicolour_redblack(scheme, order):
    _per_phase(scheme, order, "gs_multicolour", "gs_redblack")


@pytest.mark.parametrize("pres", ["mg", "cg", "jacobi"])
def test_per_phase_pressure_solvers(pres):
    _per_phase("QUICK", 4, "gs_multicolour", pres)


def test_per_phase_jacobi_momentum():
    _per_phase("QUICK", 4, "jacobi", "gs_redblack")


@pytest.mark.parametrize("pres", ["gs_redblack", "mg", "cg"])
def test_fixed_work_run_bit_identical(pres):
    """200 outer iterations (graph replays) give the oracle's fields and residual log."""
    n = 200
    p, _ = sg.params_from_config(_cfg("QUICK", 4, "gs_multicolour", pres, n))
    cpu = po.Oracle(p)
    nc, logc, convc = cpu.run()
    gpu = sg.Solver(p)
    done, logg, convg = gpu.advance(n)
    print(f"fixed work {pres}: {done} iterations", flush=True)
    assert done == n and not convg
    assert bits_equal(logg, logc[:, 1:]), first_diff(logg, logc[:, 1:])
    for f in ["u", "v", "p"]:
        _same(f, gpu.get(f), cpu.get(f))
    gpu.close()


def test_runs_are_deterministic():
    p, _ = sg.params_from_config(with_keys(golden_cfg("short_16"), maxit=50,
                                           momentum_solver="gs_multicolour",
                                           pressure_solver="mg", backend="gpu"))
    a, b = sg.Solver(p), sg.Solver(p)
    _, la, _ = a.advance(50)
    _, lb, _ = b.advance(50)
    assert bits_equal(la, lb)
    _same("u", a.get("u"), b.get("u"))


def test_stop_rule_matches_oracle_to_convergence():
    """Config 1 (32x32, Re 100, QUICK/2, multicolour + red-black, tol 1e-5) to tolerance: the
    device-side stop rule stops at the oracle's iteration with the oracle's fields."""
    cfg = ("N_cells_x = 32\nN_cells_y = 32\nRe = 100\nconvection_scheme = QUICK\n"
           "diffusion_order = 2\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = gs_redblack\nerror_tolerance = 1e-5\nmaxit = 20000\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    nc, logc, convc = cpu.run()
    gpu = sg.Solver(p)
    total, logs, conv = 0, [], False
    while total < p.maxit and not conv:
        k, lg, conv = gpu.advance(500)
        total += k
        logs.append(lg)
    print(f"config-1: oracle {nc} iterations, gpu {total}", flush=True)
    assert convc and conv and total == nc
    assert bits_equal(np.concatenate(logs), logc[:, 1:])
    for f in ["u", "v", "p"]:
        _same(f, gpu.get(f), cpu.get(f))


def test_running_max_normalisation_stop():
    cfg = with_keys(golden_cfg("converged_32x40"), momentum_solver="gs_multicolour",
                    pressure_solver="mg", maxit=3000, error_tolerance=1e-3, backend="gpu")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    nc, logc, convc = cpu.run()
    gpu = sg.Solver(p)
    r = gpu.outer_iteration()  # the host loop's first, eager iteration
    total, conv = 1, False
    while total < p.maxit and not conv:
        k, lg, conv = gpu.advance(400)
        total += k
    assert convc and conv and total == nc, (nc, total)
    _same("u", gpu.get("u"), cpu.get("u"))


@pytest.mark.parametrize("nx,ny,inner", [
    (50, 41, 1), (77, 130, 3), (96, 96, 4), (130, 128, 5), (257, 96, 9), (200, 301, 8),
])
def test_redblack_staged_tiles_bit_identical(nx, ny, inner):
    """The red-black p' solve on grids too large for one workgroup: tiles of 24 x 32 cells and a
    halo of eight in LDS, four sweeps per launch, on ragged sizes and sweep counts that are not
    multiples of four. After 3 oracle iterations, one assemble_p + solve_p on each side (p' and
    the p' operator, its derived AW, AS and AP included, bit for bit), then a 12-iteration run
    (graph replays) against the oracle's."""
    cfg = with_keys(golden_cfg("converged_32x40"), N_cells_x=nx, N_cells_y=ny,
                    momentum_solver="gs_multicolour", pressure_solver="gs_redblack",
                    maxit_inner_p=inner, maxit=3, error_tolerance=0, backend="gpu")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    for _ in range(3):
        cpu.outer_iteration()
    gpu = sg.Solver(p)
    for name in sg.FIELDS:
        gpu.set(name, cpu.get(name))
    print(f"red-black staged {nx}x{ny}, {inner} sweeps", flush=True)
    rc, rg = cpu.assemble_p(), gpu.assemble_p()
    assert _rel_close(rg, rc), (rg, rc)
    cpu.solve_p(); gpu.solve_p()
    _same("solve_p pc", gpu.get("pc"), cpu.get("pc"))
    for n in ["AEP", "AWP", "ANP", "ASP", "APP", "b"]:
        _same("assemble_p " + n, gpu.get(n), cpu.get(n))
    gpu.close()
    p, _ = sg.params_from_config(with_keys(cfg, maxit=12))
    cpu = po.Oracle(p)
    nc, logc, _ = cpu.run()
    gpu = sg.Solver(p)
    done, logg, _ = gpu.advance(12)
    assert done == 12
    assert bits_equal(logg, logc[:, 1:]), first_diff(logg, logc[:, 1:])
    for f in ["u", "v", "p", "pc"]:
        _same(f, gpu.get(f), cpu.get(f))
    gpu.close()


@pytest.mark.parametrize("nx,ny,pres,scheme,order,iters", [
    (33, 47, "mg", "QUICK", 4, 30),       # no coarsening: the whole p' solve in the LDS tail
    (64, 96, "mg", "QUICK", 4, 30),       # non-power-of-two hierarchy (coarsest 4 x 6)
    (96, 96, "mg", "UDS", 2, 30),         # coarsest 6 x 6, 2-colour momentum
    (130, 128, "gs_redblack", "QUICK", 4, 20),
    (5, 7, "mg", "QUICK", 4, 20),         # smaller than every tile
    (129, 129, "mg", "QUICK", 2, 5),      # single 129^2 level: coarse solve in global memory
    (256, 256, "mg", "QUICK", 4, 10),     # multi-level kernels above the LDS tail
])
def test_fixed_work_sizes_bit_identical(nx, ny, pres, scheme, order, iters):
    cfg = (f"N_cells_x = {nx}\nN_cells_y = {ny}\nRe = 1000\nconvection_scheme = {scheme}\n"
           f"diffusion_order = {order}\nmomentum_solver = gs_multicolour\n"
           f"pressure_solver = {pres}\nmaxit = {iters}\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    nc, logc, _ = cpu.run()
    gpu = sg.Solver(p)
    print(f"sizes {nx}x{ny} {pres} {scheme}/{order}: {iters} iterations", flush=True)
    done, logg, _ = gpu.advance(iters)
    assert done == iters
    assert bits_equal(logg, logc[:, 1:]), first_diff(logg, logc[:, 1:])
    # (After outer iterations AWP, ASP and APP may be derived from AEP and ANP on download.)
    for f in ["u", "v", "p", "pc", "b", "cu_AP", "cv_AP", "AEP", "AWP", "ANP", "ASP", "APP"]:
        _same(f, gpu.get(f), cpu.get(f))
    gpu.close()


@pytest.mark.parametrize("threads,mid,lds", [(64, 0, 256), (1024, 65536, 4096), (256, 4096, 0)])
def test_tuning_knobs_do_not_change_results(threads, mid, lds):
    cfg = ("N_cells_x = 256\nN_cells_y = 256\nRe = 1000\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = mg\nmaxit = 8\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    _, logc, _ = cpu.run()
    gpu = sg.Solver(p)
    gpu.set_tuning(sg.TUNE_TAIL_THREADS, threads)
    gpu.set_tuning(sg.TUNE_MID_CELLS, mid)
    gpu.set_tuning(sg.TUNE_LDS_CELLS, lds)
    done, logg, _ = gpu.advance(8)
    assert bits_equal(logg, logc[:, 1:]), first_diff(logg, logc[:, 1:])
    _same("p", gpu.get("p"), cpu.get("p"))


@pytest.mark.parametrize("staged", [0, 2, 64])
@pytest.mark.parametrize("pre,post", [(2, 2), (1, 3), (3, 0)])
@pytest.mark.parametrize("nx,ny", [(256, 192), (640, 528), (612, 520)])
def test_staged_multigrid_tiles_bit_identical(staged, pre, post, nx, ny):
    """Multigrid tiles that stage the operator in LDS from level `staged` on (64: none), with the
    one-workgroup tail cut to the coarsest level so every other level runs tiled. 256 x 192 makes
    partial 16 x 16 tiles on every level; 640 x 528 and 612 x 520 put level 0 on the large tiles
    (24 x 32 in FP64; partial at the edges) and the rest on 16 x 16."""
    cfg = (f"N_cells_x = {nx}\nN_cells_y = {ny}\nRe = 1000\nmomentum_solver = gs_multicolour\n"
           f"pressure_solver = mg\nmg_pre = {pre}\nmg_post = {post}\nmaxit = 6\n"
           "error_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    _, logc, _ = cpu.run()
    gpu = sg.Solver(p)
    gpu.set_tuning(sg.TUNE_MID_CELLS, 0)
    gpu.set_tuning(sg.TUNE_LDS_CELLS, 0)
    gpu.set_tuning(sg.TUNE_MG_STAGED, staged)
    print(f"{nx}x{ny} staged from {staged}, pre {pre} post {post}", flush=True)
    done, logg, _ = gpu.advance(6)
    assert bits_equal(logg, logc[:, 1:]), first_diff(logg, logc[:, 1:])
    for f in ("p", "pc"):
        _same(f, gpu.get(f), cpu.get(f))
    gpu.close()


@pytest.mark.parametrize("staged", [0, 1])
def test_staged_multigrid_tiles_mixed_bit_identical(staged):
    """The FP32 V-cycle of the mixed form gives the same bits with and without staged tiles
    (level 0 on 32 x 32 tiles, the others on 16 x 16)."""
    cfg = ("N_cells_x = 640\nN_cells_y = 512\nRe = 400\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = mg\nmg_precision = mixed\nmaxit = 8\nerror_tolerance = 0\n"
           "backend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    runs = []
    for st in (64, staged):
        gpu = sg.Solver(p)
        gpu.set_tuning(sg.TUNE_MID_CELLS, 0)
        gpu.set_tuning(sg.TUNE_MG_STAGED, st)
        _, lg, _ = gpu.advance(8)
        runs.append((lg, gpu.get("p"), gpu.get("u")))
        gpu.close()
    for a, b in zip(runs[0], runs[1]):
        assert bits_equal(a, b)


def test_autotune_restores_state_and_keeps_parity():
    cfg = ("N_cells_x = 128\nN_cells_y = 128\nRe = 1000\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = mg\nmaxit = 12\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    _, logc, _ = cpu.run()
    gpu = sg.Solver(p)
    gpu.advance(6)
    before = {n: gpu.get(n) for n in ("b", "pc", "u")}
    report = gpu.autotune(reps=3)
    print(report, flush=True)
    assert "best:" in report
    for n, a in before.items():
        _same("after autotune " + n, gpu.get(n), a)
    _, logg, _ = gpu.advance(6)
    assert bits_equal(logg, logc[6:, 1:])
    _same("u", gpu.get("u"), cpu.get("u"))


@pytest.mark.parametrize("tile", range(len(sg.MOM_TILES)))
@pytest.mark.parametrize("scheme,order,inner", [("QUICK", 4, 4), ("UDS", 2, 3), ("QUICK", 2, 1)])
def test_sweep_tile_shapes_bit_identical(tile, scheme, order, inner):
    """Every momentum sweep tile shape, odd and even sweep counts, 5 and 2 colours."""
    cfg = (f"N_cells_x = 130\nN_cells_y = 200\nRe = 1000\nconvection_scheme = {scheme}\n"
           f"diffusion_order = {order}\nmomentum_solver = gs_multicolour\n"
           f"maxit_inner_u = {inner}\nmaxit_inner_v = {inner + 1}\n"
           "pressure_solver = mg\nmaxit = 6\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    _, logc, _ = cpu.run()
    gpu = sg.Solver(p)
    gpu.set_tuning(sg.TUNE_MOM_TILE, tile)
    gpu.set_tuning(sg.TUNE_MOM_TILE_FIRST, (tile + 3) % len(sg.MOM_TILES))  # its own shape
    done, logg, _ = gpu.advance(6)
    assert bits_equal(logg, logc[:, 1:]), first_diff(logg, logc[:, 1:])
    for f in ("u", "v", "p"):
        _same(f, gpu.get(f), cpu.get(f))


@pytest.mark.parametrize("scheme,order", [("QUICK", 4), ("UDS", 2), ("UDS", 4)])
def test_face_form_after_advance(scheme, order):
    """Outer iterations keep the momentum coefficients as face velocities (multicolour sweeps):
    downloads expand them, a per-phase sweep uses them as they are, a per-phase residual and an
    uploaded coefficient switch to the explicit arrays; every result bit for bit."""
    cfg = (f"N_cells_x = 130\nN_cells_y = 200\nRe = 1000\nconvection_scheme = {scheme}\n"
           f"diffusion_order = {order}\nmomentum_solver = gs_multicolour\n"
           "maxit_inner_u = 3\nmaxit_inner_v = 4\n"
           "pressure_solver = mg\nmaxit = 3\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    for _ in range(3):
        cpu.outer_iteration()
    # Downloads after the iterations: every coefficient of both components.
    g1 = sg.Solver(p)
    g1.advance(3)
    for comp in ("cu_", "cv_"):
        for c in COEFS:
            _same("advance " + comp + c, g1.get(comp + c), cpu.get(comp + c))
    # A per-phase sweep on the face form, then a residual (explicit arrays).
    g2 = sg.Solver(p)
    g2.advance(3)
    cpu.sweep_u(2); g2.sweep_u(2)
    _same("sweep_u after advance", g2.get("u"), cpu.get("u"))
    rc, rg = cpu.residual_v(), g2.residual_v()
    assert _rel_close(rg, rc), (rg, rc)
    cpu.sweep_v(1); g2.sweep_v(1)
    _same("sweep_v after residual_v", g2.get("v"), cpu.get("v"))
    # An uploaded coefficient replaces the face form.
    ae = cpu.get("cu_AE") * 1.001
    cpu.set("cu_AE", ae); g2.set("cu_AE", ae)
    cpu.sweep_u(1); g2.sweep_u(1)
    _same("sweep_u after upload", g2.get("u"), cpu.get("u"))
    for c in COEFS:
        _same("after upload cu_" + c, g2.get("cu_" + c), cpu.get("cu_" + c))


def test_lean_pressure_assembly_after_advance():
    """With the FP64 multigrid staged from level 0 and the tail below level 1, outer iterations
    write only AE, AN and b of the p' operator; AW, AS and AP are derived on demand, bit for bit,
    for downloads and for a per-phase p' solve."""
    cfg = ("N_cells_x = 256\nN_cells_y = 256\nRe = 1000\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = mg\nmaxit = 4\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    for _ in range(4):
        cpu.outer_iteration()
    for order in ("download", "solve"):
        gpu = sg.Solver(p)
        gpu.set_tuning(sg.TUNE_MID_CELLS, 0)  # the tail from 64x64: level 0 and 1 staged
        gpu.set_tuning(sg.TUNE_MG_STAGED, 0)
        done, _, _ = gpu.advance(4)
        assert done == 4
        if order == "download":
            for f in ["AEP", "AWP", "ANP", "ASP", "APP", "b", "pc", "u", "v", "p"]:
                _same("lean " + f, gpu.get(f), cpu.get(f))
        else:
            ref = po.Oracle(p)
            for _ in range(4):
                ref.outer_iteration()
            ref.solve_p(); gpu.solve_p()
            _same("solve_p after lean iterations", gpu.get("pc"), ref.get("pc"))
        gpu.close()


def test_lean_pressure_assembly_mixed():
    """The mixed form with a lean p' assembly (its FP64 conversion and defect kernels derive AW,
    AS and AP) gives the same bits as with the full assembly (unstaged tiles, not lean)."""
    cfg = ("N_cells_x = 256\nN_cells_y = 256\nRe = 1000\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = mg\nmg_precision = mixed\nmg_dc_maxit = 2\nmaxit = 5\n"
           "error_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    runs = []
    for staged in (0, 64):  # 0: staged from level 0, lean; 64: unstaged, full assembly
        gpu = sg.Solver(p)
        gpu.set_tuning(sg.TUNE_MID_CELLS, 0)
        gpu.set_tuning(sg.TUNE_MG_STAGED, staged)
        done, log, _ = gpu.advance(5)
        runs.append((log, [gpu.get(f) for f in ("u", "v", "p", "pc", "AWP", "ASP", "APP")]))
        gpu.close()
    assert bits_equal(runs[0][0], runs[1][0]), first_diff(runs[0][0], runs[1][0])
    for a, b in zip(runs[0][1], runs[1][1]):
        assert bits_equal(a, b), first_diff(a, b)


def test_phase_timing_leaves_results_unchanged():
    """Phase timing of replayed iterations (the first iteration of every sg_advance call replayed
    with event-record nodes at the phase boundaries) gives the oracle's results bit for bit, and
    per-phase device times that add up to about the replayed iteration's time."""
    import time
    cfg = ("N_cells_x = 256\nN_cells_y = 256\nRe = 1000\nmomentum_solver = gs_multicolour\n"
           "pressure_solver = mg\nmaxit = 100\nerror_tolerance = 0\nbackend = gpu\n")
    p, _ = sg.params_from_config(cfg)
    cpu = po.Oracle(p)
    _, logc, _ = cpu.run()
    g = sg.Solver(p)
    g.set_phase_timing(True)
    logs = []
    t0 = time.perf_counter()
    for _ in range(10):
        _, lg, _ = g.advance(10)
        logs.append(lg)
    wall = time.perf_counter() - t0
    assert bits_equal(np.concatenate(logs), logc[:, 1:])
    _same("u", g.get("u"), cpu.get("u"))
    acc, n = g.phase_times()
    per_iter = acc.sum() / n
    print(f"phase-timed iterations {n}: " + ", ".join(
        f"{name} {1e6 * t / n:.1f} us" for name, t in zip(sg.PHASES, acc)) +
        f"; sum {1e6 * per_iter:.1f} us, wall {1e6 * wall / 100:.1f} us per iteration",
        flush=True)
    assert n == 10 and np.all(acc[[0, 3, 4, 7, 8, 9, 10]] > 0)
    assert 0.5 * wall / 100 < per_iter < 1.5 * wall / 100
    g.close()
