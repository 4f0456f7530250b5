"""The oracle's FP32 momentum defect correction (momentum_precision = mixed; the GPU form, with no
reference counterpart), pinned on the CPU by what the mathematics fixes:

 - a multicolour sweep is linear in the field, so the correction form reproduces the reference's
   FP64 sweeps (src/momentum_u.cpp:123-143) up to FP32 rounding of the correction: after n
   sweeps from a realistic state, x_dc - x0 equals x_fp64 - x0 to a few FP32 ulps of the largest
   change (a dropped term, a wrong sign or a wrong neighbour gives an O(1) relative difference);
 - r0 = 0 is a fixed point: a state whose residual is exactly zero does not move;
 - the correction is linear in r0: scaling the pressure term's difference scales nothing else;
 - whole runs (config 1: 32 x 32, Re 100, QUICK / 2nd order, multicolour + red-black, to 1e-5)
   converge in the FP64 run's iteration count within 2 %, to fields within 1e-9 of it.
"""
import numpy as np
import pytest

import pyoracle as po
import simple_gpu as sg


def _params(n=40, m=52, re=400, scheme="QUICK", order=4, **kv):
    text = (f"N_cells_x = {n}\nN_cells_y = {m}\nRe = {re}\nconvection_scheme = {scheme}\n"
            f"diffusion_order = {order}\nmomentum_solver = gs_multicolour\n"
            f"pressure_solver = mg\nmaxit = 100000\nerror_tolerance = 0\n"
            + "".join(f"{k} = {v}\n" for k, v in kv.items()))
    return sg.params_from_config(text)[0]


@pytest.mark.parametrize("scheme,order", [("QUICK", 4), ("UDS", 2), ("QUICK", 2)])
@pytest.mark.parametrize("comp", ["u", "v"])
@pytest.mark.parametrize("nsweeps", [1, 4])
def test_correction_form_equals_fp64_sweeps(scheme, order, comp, nsweeps):
    p = _params(scheme=scheme, order=order)
    base = po.Oracle(p)
    for _ in range(6):
        base.outer_iteration()
    state = {f: base.get(f) for f in po.FIELDS}

    def fresh():
        o = po.Oracle(p)
        for f, a in state.items():
            o.set(f, a)
        getattr(o, "assemble_" + comp)()
        getattr(o, "bc_" + comp)()
        return o

    a, b = fresh(), fresh()
    x0 = a.get(comp)
    getattr(a, "sweep_" + comp)(nsweeps)         # FP64, the reference's sweeps
    getattr(b, "sweep_" + comp + "_dc")(nsweeps)  # FP32 correction
    xa, xb = a.get(comp), b.get(comp)
    change = np.abs(xa - x0).max()
    err = np.abs(xb - xa).max()
    print(f"{scheme}/{order} {comp} x{nsweeps}: max change {change:.3e}, max |dc - fp64| "
          f"{err:.3e} ({err / change:.2e} of the change)", flush=True)
    assert change > 0
    assert err <= 2e-6 * change
    # outside the sweep range nothing moves
    outside = np.ones_like(xa, dtype=bool)
    i0, i1 = (p.H, p.NX + p.H - 1) if comp == "u" else (p.H + 1, p.NX + p.H - 1)
    j0, j1 = (p.H + 1, p.NY + p.H - 1) if comp == "u" else (p.H + 1, p.NY + p.H)
    outside[i0:i1 + 1, j0:j1 + 1] = False
    assert np.array_equal(xb[outside], x0[outside])


def test_zero_residual_is_a_fixed_point():
    p = _params(n=24, m=24)
    o = po.Oracle(p)
    for _ in range(4):
        o.outer_iteration()
    o.assemble_u()
    o.bc_u()
    # Make x0 an exact solution of the cell equations: AP x = nb(x) + pterm with pterm chosen
    # per cell, i.e. replace p so that the residual vanishes is awkward; instead zero the field,
    # the pressure and the coefficients' pressure term source: x = 0, p = 0 gives r0 = 0.
    o.set("u", np.zeros((p.nrows, p.ncols)))
    o.set("p", np.zeros((p.nrows, p.ncols)))
    o.sweep_u_dc(4)
    assert np.all(o.get("u") == 0.0)


def test_correction_is_linear_in_the_residual():
    p = _params(n=24, m=30)
    o = po.Oracle(p)
    for _ in range(5):
        o.outer_iteration()
    st = {f: o.get(f) for f in po.FIELDS}
    outs = []
    for scale in (1.0, 2.0):
        q = po.Oracle(p)
        for f, a in st.items():
            q.set(f, a)
        q.assemble_u()
        q.bc_u()
        q.set("u", np.zeros((p.nrows, p.ncols)))  # x0 = 0: r0 = pterm only
        q.set("p", st["p"] * scale)
        q.sweep_u_dc(4)
        outs.append(q.get("u"))
    # doubling p doubles every pressure term (a difference times dy, exact) and so r0 and d
    assert np.array_equal(outs[1], 2.0 * outs[0])


def test_whole_run_config1_matches_fp64():
    base = ("N_cells_x = 32\nN_cells_y = 32\nRe = 100\nconvection_scheme = QUICK\n"
            "diffusion_order = 2\nmomentum_solver = gs_multicolour\n"
            "pressure_solver = gs_redblack\nerror_tolerance = 1e-5\nmaxit = 20000\n")
    p64 = sg.params_from_config(base)[0]
    pdc = sg.params_from_config(base + "momentum_precision = mixed\n")[0]
    a, b = po.Oracle(p64), po.Oracle(pdc)
    na, _, ca = a.run()
    nb, _, cb = b.run()
    errs = [float(np.abs(a.get(f) - b.get(f)).max()) for f in ("u", "v", "p")]
    print(f"config 1: fp64 {na} iterations, mixed momentum {nb}; max |du| {errs[0]:.2e} "
          f"|dv| {errs[1]:.2e} |dp| {errs[2]:.2e}", flush=True)
    assert ca and cb and abs(nb - na) <= 0.02 * na
    assert max(errs) < 1e-9
