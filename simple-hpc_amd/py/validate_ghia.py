"""Centreline errors of a run against Ghia, Ghia and Shin (1982), Re 400 and 1000.

Restates the reference's validation (postprocess/validate_ghia.py:31-99 with
postprocess/fields.py:6-34, postprocess/ghia.py:40-70 and postprocess/io.py:11-44): the run's
cell-centred velocities give the vertical centreline of u and the horizontal centreline of v
(the two middle columns or rows averaged when the count is even, walls added: u = 0 at y = 0 and
1 at the lid, v = 0 at both walls); each is interpolated linearly at Ghia's 15 interior points,
and the RMS and largest absolute error are reported. Ghia's data (ghia1982.json) are the values
of the published tables, with the same excluded point (Re 400, v at x = 0.90625, a likely
misprint, kept with --keep-suspect). Field files may be .csv (the reference's format) or .npy
(field_format = npy).

Usage: validate_ghia.py RUN_DIR ... [--markdown] [--keep-suspect]
                        [--check --max-rms-u X --max-rms-v Y]
With --check the exit status is 1 if any run exceeds the limits (the reference's CTest limits,
CMakeLists.txt:139-150: 0.010 / 0.010 at Re 400 and 0.025 / 0.015 at Re 1000).
"""
import argparse
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = ["name", "Re", "NX", "iterations", "converged", "rms_u", "max_u", "rms_v", "max_v"]
# The reference's CTest limits (CMakeLists.txt:139-150): Re -> (max rms_u, max rms_v).
LIMITS = {400: (0.010, 0.010), 1000: (0.025, 0.015)}

_DATA = None


def _data():
    global _DATA
    if _DATA is None:
        with open(os.path.join(HERE, "ghia1982.json")) as f:
            _DATA = json.load(f)
    return _DATA


def _coords(index):
    return (np.asarray(index, dtype=float) - 1.0) / (_data()["grid"] - 1)


def ghia_u(re):
    """(y, u) on the vertical centreline, y increasing."""
    d = _data()
    if str(re) not in d["u"]:
        raise ValueError(f"Ghia data available for Re 400 and 1000, got {re}")
    y = _coords(d["u_y_index"])
    o = np.argsort(y)
    return y[o], np.asarray(d["u"][str(re)], dtype=float)[o]


def ghia_v(re):
    """(x, v) on the horizontal centreline, x increasing."""
    d = _data()
    if str(re) not in d["v"]:
        raise ValueError(f"Ghia data available for Re 400 and 1000, got {re}")
    x = _coords(d["v_x_index"])
    o = np.argsort(x)
    return x[o], np.asarray(d["v"][str(re)], dtype=float)[o]


def suspect(re, comp):
    return _data()["suspect"].get(str(re), {}).get(comp, [])


def load_field(path, name):
    """NAME.csv, else NAME.npy (plain loads: no pickled objects), else None."""
    csv = os.path.join(path, name + ".csv")
    if os.path.isfile(csv):
        return np.loadtxt(csv, delimiter=",", ndmin=2)
    npy = os.path.join(path, name + ".npy")
    if os.path.isfile(npy):
        return np.load(npy, allow_pickle=False)
    return None


def _mid(a, axis):
    n = a.shape[axis]
    if n % 2 == 0:
        return 0.5 * (np.take(a, n // 2 - 1, axis=axis) + np.take(a, n // 2, axis=axis))
    return np.take(a, n // 2, axis=axis)


def centrelines(ux, uy):
    """Centrelines from the cell-centred velocities (row 0 is the lid row): dict(y, u, x, v),
    walls included, coordinates increasing."""
    nx, ny = ux.shape
    x = (np.arange(ny) + 0.5) / ny
    y = (np.arange(nx) + 0.5) / nx
    u_col = _mid(ux, 1)[::-1]  # the lid row first in the file: flip so y increases
    v_row = _mid(uy, 0)
    return dict(y=np.concatenate(([0.0], y, [1.0])), u=np.concatenate(([0.0], u_col, [1.0])),
                x=np.concatenate(([0.0], x, [1.0])), v=np.concatenate(([0.0], v_row, [0.0])))


def errors(coord, vals, gcoord, gvals, skip=()):
    """RMS and largest absolute error of the interpolated centreline at Ghia's interior points
    (the wall points dropped, points within 1e-12 of `skip` excluded)."""
    gc, gv = gcoord[1:-1], gvals[1:-1]
    keep = np.array([not any(abs(x - s) < 1e-12 for s in skip) for x in gc], dtype=bool)
    err = (np.interp(gc, coord, vals) - gv)[keep]
    return float(np.sqrt(np.mean(err ** 2))), float(np.max(np.abs(err)))


def evaluate(path, keep_suspect=False):
    with open(os.path.join(path, "run_info.json")) as f:
        info = json.load(f)
    re = int(round(float(info["Re"])))
    ux, uy = load_field(path, "ux_centre"), load_field(path, "uy_centre")
    if ux is None or uy is None:
        raise ValueError(f"{path}: no ux_centre / uy_centre field (write_fields = False?)")
    c = centrelines(ux, uy)
    gy, gu = ghia_u(re)
    gx, gv = ghia_v(re)
    rms_u, max_u = errors(c["y"], c["u"], gy, gu, [] if keep_suspect else suspect(re, "u"))
    rms_v, max_v = errors(c["x"], c["v"], gx, gv, [] if keep_suspect else suspect(re, "v"))
    return [os.path.basename(os.path.normpath(path)), re, int(info["N_cells_x"]),
            int(info["iterations"]), "yes" if info["converged"] else "no",
            rms_u, max_u, rms_v, max_v]


def _fmt(row):
    return [f"{x:.5f}" if isinstance(x, float) else str(x) for x in row]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("runs", nargs="+")
    ap.add_argument("--markdown", action="store_true")
    ap.add_argument("--keep-suspect", action="store_true", help="include suspect Ghia points")
    ap.add_argument("--check", action="store_true", help="exit 1 if an rms limit is exceeded")
    ap.add_argument("--max-rms-u", type=float, default=None)
    ap.add_argument("--max-rms-v", type=float, default=None)
    a = ap.parse_args(argv)
    if a.check and (a.max_rms_u is None or a.max_rms_v is None):
        ap.error("--check requires --max-rms-u and --max-rms-v")
    rows = [evaluate(r, a.keep_suspect) for r in a.runs
            if os.path.isfile(os.path.join(r, "run_info.json"))]
    if a.check and not rows:
        print("no runs found")
        return 1
    if not a.keep_suspect:
        print(f"excluded suspect Ghia points: {_data()['suspect']} (use --keep-suspect to include)")
    if a.markdown:
        print("| " + " | ".join(HEADER) + " |")
        print("|" + "|".join("---" for _ in HEADER) + "|")
        for row in rows:
            print("| " + " | ".join(_fmt(row)) + " |")
    else:
        for row in rows:
            print("  ".join(_fmt(row)))
    status = 0
    if a.check:
        for row in rows:
            ru, rv = row[HEADER.index("rms_u")], row[HEADER.index("rms_v")]
            if ru > a.max_rms_u or rv > a.max_rms_v:
                print(f"FAIL {row[0]}: rms_u {ru:.5f} (max {a.max_rms_u}), "
                      f"rms_v {rv:.5f} (max {a.max_rms_v})")
                status = 1
    return status


if __name__ == "__main__":
    sys.exit(main())
