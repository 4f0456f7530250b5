"""Brute-force known answers for the inner solvers (CPU), restating the reference's
tests/unit/test_inner_solvers.cpp:60-190: random coefficients from its LCG (seed 12345) on a 9x11
grid, one sweep of the oracle against a naive loop written here, bit for bit.
"""
import numpy as np
import pytest

import pyoracle as po
import simple_gpu as sg


class Lcg:
    """The reference's deterministic generator (test_inner_solvers.cpp:22-28)."""

    def __init__(self, seed=12345):
        self.s = seed

    def next(self, lo):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        return lo + float(self.s >> 11) / 9007199254740992.0

    def fill(self, shape, lo):
        a = np.empty(shape)
        for i in range(shape[0]):
            for j in range(shape[1]):
                a[i, j] = self.next(lo)
        return a


def small(scheme, order, **kv):
    text = f"N_cells_x = 9\nN_cells_y = 11\nRe = 100\nconvection_scheme = {scheme}\n" \
           f"diffusion_order = {order}\n" + "".join(f"{k} = {v}\n" for k, v in kv.items())
    return sg.params_from_config(text)[0]


NB = ["AE", "AW", "AN", "AS", "AEE", "AWW", "ANN", "ASS"]


def random_system(p, g, prefix):
    shape = (p.nrows, p.ncols)
    c = {n: g.fill(shape, -0.5) for n in NB}
    c["AP"] = g.fill(shape, 4.0)
    return c


def nb9(c, f, i, j):
    return ((c["AE"][i, j] * f[i, j + 1] + c["AW"][i, j] * f[i, j - 1] + c["AN"][i, j] * f[i - 1, j]
             + c["AS"][i, j] * f[i + 1, j])
            + (c["AEE"][i, j] * f[i, j + 2] + c["AWW"][i, j] * f[i, j - 2]
               + c["ANN"][i, j] * f[i - 2, j] + c["ASS"][i, j] * f[i + 2, j]))


def nb5(c, f, i, j):
    return (c["AE"][i, j] * f[i, j + 1] + c["AW"][i, j] * f[i, j - 1] + c["AN"][i, j] * f[i - 1, j]
            + c["AS"][i, j] * f[i + 1, j])


def load(o, prefix, c):
    for n, a in c.items():
        o.set(prefix + n, a)


def test_colourings_separate_all_stencil_offsets():
    far9 = [(0, 1), (0, -1), (0, 2), (0, -2), (1, 0), (-1, 0), (2, 0), (-2, 0)]
    near5 = [(0, 1), (0, -1), (1, 0), (-1, 0)]
    col = {5: lambda i, j: (i + 3 * j) % 5, 2: lambda i, j: (i + j) % 2}
    for ncol, offs in ((5, far9), (2, near5)):
        for i in range(2, 22):
            for j in range(2, 22):
                for di, dj in offs:
                    assert col[ncol](i, j) != col[ncol](i + di, j + dj)


@pytest.mark.parametrize("comp", ["u", "v"])
def test_jacobi_sweep_equals_reference_loop(comp):
    p = small("QUICK", 4)
    g = Lcg()
    c = random_system(p, g, "")
    x = g.fill((p.nrows, p.ncols), -0.5)
    pf = g.fill((p.nrows, p.ncols), -0.5)
    o = po.Oracle(p)
    load(o, "cu_" if comp == "u" else "cv_", c)
    o.set(comp, x)
    o.set("p", pf)
    ref = x.copy()
    H, NX, NY = p.H, p.NX, p.NY
    if comp == "u":
        a = p.relaxation_factor_u
        for i in range(NX + H - 1, H - 1, -1):
            for j in range(H + 1, NY + H):
                ref[i, j] = a * (nb9(c, x, i, j) + (pf[i, j] - pf[i, j + 1]) * p.dy) / c["AP"][i, j] \
                    + (1 - a) * x[i, j]
        o.sweep_u_kind(sg.MOM_JACOBI, 1)
    else:
        a = p.relaxation_factor_v
        for i in range(NX + H - 1, H, -1):
            for j in range(H + 1, NY + H + 1):
                ref[i, j] = a * (nb9(c, x, i, j) + (pf[i, j] - pf[i - 1, j]) * p.dx) / c["AP"][i, j] \
                    + (1 - a) * x[i, j]
        o.sweep_v_kind(sg.MOM_JACOBI, 1)
    assert np.array_equal(o.get(comp), ref)


@pytest.mark.parametrize("scheme,order", [("QUICK", 4), ("UDS", 2)])
@pytest.mark.parametrize("comp", ["u", "v"])
def test_multicolour_sweep_equals_colour_by_colour_loop(scheme, order, comp):
    p = small(scheme, order)
    ncol = 2 if (scheme == "UDS" and order == 2) else 5
    g = Lcg()
    c = random_system(p, g, "")
    if ncol == 2:
        for n in ("AEE", "AWW", "ANN", "ASS"):
            c[n][:] = 0.0
    x = g.fill((p.nrows, p.ncols), -0.5)
    pf = g.fill((p.nrows, p.ncols), -0.5)
    o = po.Oracle(p)
    load(o, "cu_" if comp == "u" else "cv_", c)
    o.set(comp, x)
    o.set("p", pf)
    ref = x.copy()
    H, NX, NY = p.H, p.NX, p.NY
    colour = (lambda i, j: (i + j) % 2) if ncol == 2 else (lambda i, j: (i + 3 * j) % 5)
    for col in range(ncol):
        if comp == "u":
            rows, cols, a = range(NX + H - 1, H - 1, -1), range(H + 1, NY + H), p.relaxation_factor_u
        else:
            rows, cols, a = range(NX + H - 1, H, -1), range(H + 1, NY + H + 1), p.relaxation_factor_v
        for i in rows:
            for j in cols:
                if colour(i, j) != col:
                    continue
                s = nb5(c, ref, i, j) if ncol == 2 else nb9(c, ref, i, j)
                pt = (pf[i, j] - pf[i, j + 1]) * p.dy if comp == "u" else (pf[i, j] - pf[i - 1, j]) * p.dx
                ref[i, j] = a * (s + pt) / c["AP"][i, j] + (1 - a) * ref[i, j]
    if comp == "u":
        o.sweep_u_kind(sg.MOM_GS_MULTICOLOUR, 1)
    else:
        o.sweep_v_kind(sg.MOM_GS_MULTICOLOUR, 1)
    assert np.array_equal(o.get(comp), ref)


def test_pressure_damped_jacobi_and_redblack_equal_reference_loops():
    p = small("QUICK", 4, maxit_inner_p=3)
    g = Lcg()
    shape = (p.nrows, p.ncols)
    w = {n: g.fill(shape, 0.0) for n in ("AEP", "AWP", "ANP", "ASP", "b")}
    w["APP"] = w["AEP"] + w["AWP"] + w["ANP"] + w["ASP"]
    H, NX, NY = p.H, p.NX, p.NY

    def upd(pc, src, i, j):
        t = w["b"][i, j] / w["APP"][i, j]
        t += (w["AEP"][i, j] * src[i, j + 1] + w["AWP"][i, j] * src[i, j - 1]
              + w["ASP"][i, j] * src[i + 1, j] + w["ANP"][i, j] * src[i - 1, j]) / w["APP"][i, j]
        pc[i, j] = t

    ref = np.zeros(shape)
    for _ in range(p.maxit_inner_p):
        old = ref.copy()
        for i in range(NX + H - 1, H - 1, -1):
            for j in range(H + 1, NY + H + 1):
                upd(ref, old, i, j)
                ref[i, j] = 0.8 * ref[i, j] + (1 - 0.8) * old[i, j]
    o = po.Oracle(p)
    for n, a in w.items():
        o.set(n, a)
    o.set("pc", np.full(shape, 7.0))  # the solver must reset p' to 0
    o.solve_p_kind(sg.P_JACOBI)
    assert np.array_equal(o.get("pc"), ref)

    ref = np.zeros(shape)
    for _ in range(p.maxit_inner_p):
        for col in range(2):
            for i in range(NX + H - 1, H - 1, -1):
                for j in range(H + 1, NY + H + 1):
                    if (i + j) % 2 == col:
                        upd(ref, ref, i, j)
    o.set("pc", np.full(shape, 7.0))
    o.solve_p_kind(sg.P_GS_REDBLACK)
    assert np.array_equal(o.get("pc"), ref)
