"""FP32 momentum smoothing inside FP64 defect correction (momentum_precision = mixed) on the GPU.

 - Bit for bit against the oracle's restatement of the same form (FP64 assembly and residual,
   FP32 colour-by-colour sweeps of the correction, x0 + d in FP64; the oracle's form is pinned
   against the reference's FP64 sweeps in tests/test_momentum_dc.py): fixed-work runs over sizes
   that cut the 64 x 64 sweep tiles unevenly, both colourings, and the p' solvers.
 - Against the FP64 oracle (the reference's arithmetic): a run to tolerance takes the FP64 run's
   iteration count within 2 % (measured: the same count) and ends within 5e-10 of its fields.
 - After such outer iterations the FP64 momentum coefficients are not kept: downloading them or a
   per-phase residual says so; an assembly brings them back.
 - A batch of cavities in this form gives every cavity its solo result.
"""
import numpy as np
import pytest

from conftest import bits_equal, first_diff

pytestmark = pytest.mark.gpu

import pyoracle as po  # noqa: E402
import simple_gpu as sg  # noqa: E402


def _cfg(nx, ny, re=1000, scheme="QUICK", order=4, pres="mg", maxit=100000, tol=0, **kv):
    text = (f"N_cells_x = {nx}\nN_cells_y = {ny}\nRe = {re}\nconvection_scheme = {scheme}\n"
            f"diffusion_order = {order}\nmomentum_solver = gs_multicolour\n"
            f"pressure_solver = {pres}\nmaxit = {maxit}\nerror_tolerance = {tol}\nbackend = gpu\n"
            f"momentum_precision = mixed\n" + "".join(f"{k} = {v}\n" for k, v in kv.items()))
    return sg.params_from_config(text)[0]


def _same(name, g, c):
    assert bits_equal(g, c), f"{name}: {first_diff(g, c)}"


@pytest.mark.parametrize("nx,ny,scheme,order,pres,iters", [
    (130, 200, "QUICK", 4, "mg", 12),
    (96, 96, "UDS", 2, "mg", 12),        # two colours
    (64, 96, "QUICK", 2, "gs_redblack", 12),
    (5, 7, "QUICK", 4, "mg", 12),        # smaller than one tile
    (256, 256, "QUICK", 4, "mg", 10),
    (200, 130, "UDS", 4, "jacobi", 9),
])
def test_dc_bit_identical_to_oracle(nx, ny, scheme, order, pres, iters):
    p = _cfg(nx, ny, scheme=scheme, order=order, pres=pres, maxit=iters)
    cpu = po.Oracle(p)
    logc = np.array([cpu.outer_iteration() for _ in range(iters)])
    g = sg.Solver(p)
    print(f"dc {nx}x{ny} {scheme}/{order} {pres}: {iters} iterations", flush=True)
    done, logg, _ = g.advance(iters)
    assert done == iters
    assert bits_equal(logg, logc), first_diff(logg, logc)
    for f in ("u", "v", "p", "pc", "b"):
        _same(f, g.get(f), cpu.get(f))
    # the eager outer iteration (per-phase timers) takes the same path
    r = g.outer_iteration()
    rc = cpu.outer_iteration()
    assert bits_equal(r, rc), (r, rc)
    _same("u after eager iteration", g.get("u"), cpu.get("u"))
    g.close()


def test_dc_run_to_tolerance_close_to_fp64():
    n, tol = 64, 1e-5
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = 1000\nmomentum_solver = gs_multicolour\n"
            f"pressure_solver = mg\nerror_tolerance = {tol}\nmaxit = 20000\nbackend = gpu\n")
    p64 = sg.params_from_config(text)[0]
    cpu = po.Oracle(p64)
    n64, _, conv64 = cpu.run()
    pm = sg.params_from_config(text + "momentum_precision = mixed\nmg_precision = mixed\n")[0]
    g = sg.Solver(pm)
    total, conv = 0, False
    while total < pm.maxit and not conv:
        k, _, conv = g.advance(1000)
        total += k
    errs = [float(np.abs(g.get(f) - cpu.get(f)).max()) for f in ("u", "v")]
    H = pm.H
    pg = g.get("p")[H:n + H, H + 1:n + H + 1]
    pc = cpu.get("p")[H:n + H, H + 1:n + H + 1]
    errs.append(float(np.abs((pg - pg.mean()) - (pc - pc.mean())).max()))
    print(f"\n64^2 mixed momentum + mixed MG: {total} iterations vs fp64 {n64}; max|du| "
          f"{errs[0]:.2e} max|dv| {errs[1]:.2e} max|dp| {errs[2]:.2e}", flush=True)
    assert conv64 and conv and abs(total - n64) <= 0.02 * n64
    assert max(errs) < 5e-10  # measured: 5.0e-11, 5.1e-11, 2.3e-11
    g.close()


def test_coefficients_not_kept_after_dc_iterations():
    p = _cfg(40, 48, maxit=5)
    g = sg.Solver(p)
    g.advance(3)
    with pytest.raises(sg.GpuError, match="not kept with momentum_precision = mixed"):
        g.get("cu_AE")
    with pytest.raises(sg.GpuError, match="call sg_assemble_v first"):
        g.residual_v()
    # an assembly brings the explicit coefficients back, bit for bit as the oracle's
    cpu = po.Oracle(p)
    for _ in range(3):
        cpu.outer_iteration()
    g.assemble_u(); cpu.assemble_u()
    _same("cu_AE after assemble_u", g.get("cu_AE"), cpu.get("cu_AE"))
    g.bc_u(); cpu.bc_u()
    rg, rc = g.residual_u(), cpu.residual_u()
    assert rg == rc
    g.close()


def test_dc_batch_matches_solo():
    cases = [(16, 150.0), (24, 2900.0), (40, 1200.0), (64, 400.0)]
    plist = [_cfg(n, n, re=re, order=2, maxit=100000) for n, re in cases]
    b = sg.Batch(plist)
    its, _, _ = b.advance(25)
    assert its == [25] * len(cases)
    for k, p in enumerate(plist):
        s = sg.Solver(p)
        s.advance(25)
        c = b.cavity(k)
        for f in ("u", "v", "p"):
            _same(f"cavity {k} {f}", c.get(f), s.get(f))
        s.close()
    b.close()
