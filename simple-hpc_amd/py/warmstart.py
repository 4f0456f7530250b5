"""Training data, training and weights files of the learned p' warm start.

Data: batched cavity runs (one persistent launch per phase, simple_gpu.Batch) are advanced through
their convergence history; at chosen outer iterations every cavity contributes one pair:
  input  = the network features of its current p' system on multigrid level L (sg_warm_features:
           block sums of b, block means of a_P, log10(Re/1000), normalised), and
  target = the block means of the converged p' of the same system, divided by the same scale.
The converged p' comes from a teacher solver of the same size (FP64 multigrid, 30 V-cycles) into
which the system's arrays (AE, AW, AN, AS, AP of p' and b) are uploaded.

Training: Adam on mean squared error, minibatches of equal-sized coarse grids, run on the GPU by
the library's own bfloat16 MFMA forward and backward kernels (sg_unet_train_step).

Weights file: the ASCII line "SGUNET1 <n>\\n" then n little-endian float32 parameters.
"""
import math
import time

import numpy as np

import simple_gpu as sg

P_SYSTEM = ("AEP", "AWP", "ANP", "ASP", "APP", "b")

# (C_in as stored, C_out, C_in used) of the network's 14 3x3 convolutions, as in sg_unet.hip.
CONVS = [(8, 16, 3), (16, 16, 16), (16, 32, 16), (32, 32, 32), (32, 64, 32), (64, 64, 64),
         (64, 128, 64), (128, 128, 128), (192, 64, 192), (64, 64, 64), (96, 32, 96), (32, 32, 32),
         (48, 16, 48), (16, 16, 16)]


def init_params(seed=0):
    """He-normal convolution weights, zero biases, head weights N(0, 1/16): the flat FP32 layout of
    sg_unet_set_params, from numpy's PCG64(seed)."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, co, ci in CONVS:
        parts.append((rng.standard_normal(co * ci * 9) * np.sqrt(2.0 / (ci * 9))).astype(np.float32))
        parts.append(np.zeros(co, np.float32))
    parts.append((rng.standard_normal(16) * 0.25).astype(np.float32))
    parts.append(np.zeros(1, np.float32))
    return np.concatenate(parts)


def teacher_params(n, re):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = {re!r}\nconvection_scheme = QUICK\n"
            f"diffusion_order = 2\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_cycles = 30\nbackend = gpu\nwrite_fields = False\n")
    return sg.params_from_config(text)[0]


def block_mean(a, bs):
    nx, ny = a.shape
    return a.reshape(nx // bs, bs, ny // bs, bs).mean(axis=(1, 3))


def training_cavities(n=32, seed=1, sizes=(128, 192, 256, 384, 512), re_lo=100.0, re_hi=3200.0):
    """(N, Re) pairs for training data: N cycling over sizes, Re log-uniform, numpy PCG64(seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        u = rng.random()
        out.append((sizes[k % len(sizes)],
                    10 ** (math.log10(re_lo) + u * (math.log10(re_hi) - math.log10(re_lo)))))
    return out


def generate(cases, snapshots, level=2, precision="mixed", log=print):
    """Pairs from batched runs of `cases` sampled after each outer iteration count in `snapshots`
    (ascending). Returns {coarse shape: (x [m,h,w,8], t [m,h,w])}."""
    plist = [sg.params_from_config(sg.batch_config(n, re, precision=precision, tol=0.0))[0]
             for n, re in cases]
    b = sg.Batch(plist)
    teachers = {}
    data = {}
    bs = 1 << level
    done = 0
    t0 = time.perf_counter()
    for snap in snapshots:
        b.advance(snap - done)
        done = snap
        for k, (n, re) in enumerate(cases):
            c = b.cavity(k)
            x, scale = sg.warm_features(c, level)
            if n not in teachers:
                teachers[n] = sg.Solver(teacher_params(n, re))
            tch = teachers[n]
            for name in P_SYSTEM:
                tch.set(name, c.get(name))
            tch.solve_p()
            H = c.p.H
            pc = tch.get("pc")[H:H + c.p.NX, H + 1:H + 1 + c.p.NY]
            t = (block_mean(pc, bs) / scale).astype(np.float32)
            key = x.shape[:2]
            xs, ts = data.setdefault(key, ([], []))
            xs.append(x)
            ts.append(t)
        log(f"  data: iteration {snap}, {sum(len(v[0]) for v in data.values())} pairs, "
            f"{time.perf_counter() - t0:.1f} s")
    b.close()
    for t in teachers.values():
        t.close()
    return {k: (np.stack(v[0]), np.stack(v[1])) for k, v in data.items()}


def dihedral(a, k):
    """Image k (0..7) of the dihedral group of the square applied to the grid axes (1, 2) of a
    batch [m, h, w, ...]: k & 3 quarter turns, then a flip if k & 4. A p' system on a square grid
    reflected or turned is again a p' system (links, a_P and b carried along; a_P, the sum of the
    links, is unchanged), and its solution is the solution carried along, so each image of a pair
    is an exact pair of the image system."""
    a = np.rot90(a, k & 3, axes=(1, 2))
    if k & 4:
        a = a[:, :, ::-1]
    return np.ascontiguousarray(a)


def train(net, data, steps=2000, batch=8, lr=1e-3, seed=0, log=print, every=200, augment=False):
    """Adam steps over minibatches drawn from the size groups in proportion to their pairs; with
    augment, each minibatch is one random dihedral image of its pairs (square grids)."""
    rng = np.random.default_rng(seed)
    keys = sorted(data)
    weights = np.array([len(data[k][0]) for k in keys], np.float64)
    weights /= weights.sum()
    hist = []
    t0 = time.perf_counter()
    for s in range(steps):
        k = keys[rng.choice(len(keys), p=weights)]
        x, t = data[k]
        idx = rng.choice(len(x), size=min(batch, len(x)), replace=False)
        xb, tb = x[idx], t[idx]
        if augment and xb.shape[1] == xb.shape[2]:
            k = int(rng.integers(8))
            xb, tb = dihedral(xb, k), dihedral(tb, k)
        # cosine decay of the step size
        a = lr * 0.5 * (1 + math.cos(math.pi * s / steps))
        loss = net.train_step(xb, tb, max(a, lr * 0.02))
        hist.append(loss)
        if log and (s % every == 0 or s == steps - 1):
            log(f"  train step {s}: loss {np.mean(hist[-every:]):.4f} "
                f"({time.perf_counter() - t0:.1f} s)")
    return hist


def normalise(data):
    """Each target divided by its own RMS (all-zero targets left as they are). The solver scales
    the network's guess by the energy-minimising factor b.g / g.A.g, so only the shape of the
    prediction matters: unit-RMS targets keep the samples of very different magnitude (early and
    late in a run, small and large cavities) equally weighted in the loss."""
    out = {}
    for k, (x, t) in data.items():
        r = np.sqrt((t.astype(np.float64) ** 2).mean(axis=(1, 2)))
        r[r == 0] = 1.0
        out[k] = (x, (t / r[:, None, None]).astype(np.float32))
    return out


def shape_error(net, data):
    """Mean over samples of the relative error of the best-scaled prediction, min over a of
    |a y - t| / |t| = sqrt(1 - cos^2(y, t)) (1 for the zero prediction)."""
    errs = []
    for k, (x, t) in data.items():
        for i in range(0, len(x), 16):
            y = net.forward(x[i:i + 16]).astype(np.float64)
            tt = t[i:i + 16].astype(np.float64)
            for a, b in zip(y, tt):
                na, nb = np.linalg.norm(a), np.linalg.norm(b)
                c = float(np.dot(a.ravel(), b.ravel()) / (na * nb)) if na > 0 and nb > 0 else 0.0
                errs.append(np.sqrt(max(0.0, 1.0 - c * c)))
    return float(np.mean(errs))


def evaluate(net, data):
    """Mean squared error of the network and of the zero prediction on data."""
    num, den, cnt = 0.0, 0.0, 0
    for k, (x, t) in data.items():
        for i in range(0, len(x), 16):
            y = net.forward(x[i:i + 16])
            num += float(((y - t[i:i + 16]) ** 2).sum())
            den += float((t[i:i + 16] ** 2).sum())
            cnt += t[i:i + 16].size
    return num / cnt, den / cnt


def save(path, theta):
    theta = np.ascontiguousarray(theta, dtype="<f4")
    with open(path, "wb") as f:
        f.write(f"SGUNET1 {theta.size}\n".encode())
        f.write(theta.tobytes())


def load(path):
    with open(path, "rb") as f:
        head = f.readline().decode().split()
        if len(head) != 2 or head[0] != "SGUNET1":
            raise ValueError(f"{path}: not a warm-start weights file")
        n = int(head[1])
        theta = np.frombuffer(f.read(4 * n), dtype="<f4")
    if theta.size != n:
        raise ValueError(f"{path}: truncated")
    return theta.astype(np.float32)


def _mixed_params(n, re, learned, cycles=1, tol=1e-5, check_every=200):
    text = (f"N_cells_x = {n}\nN_cells_y = {n}\nRe = {re!r}\nconvection_scheme = QUICK\n"
            f"diffusion_order = 4\nmomentum_solver = gs_multicolour\npressure_solver = mg\n"
            f"mg_pre = 2\nmg_post = 2\nmg_cycles = {cycles}\nmg_precision = mixed\n"
            f"error_tolerance = {tol!r}\nmaxit = 10000000\nbackend = gpu\nwrite_fields = False\n"
            f"check_every = {check_every}\n")
    if learned:
        text += "pressure_warm_start = learned\n"
    return sg.params_from_config(text)[0]


def ab_per_solve(net, n, re, snapshots, level=-1):
    """Error of p' after one mixed-precision p' solve, started from zero and from the network's
    guess, against the converged p' of the same system, on states of an n x n run at Re.
    Returns a list of (iteration, error zero start, error warm start, guess kept)."""
    run = sg.Solver(_mixed_params(n, re, False))
    zero = sg.Solver(_mixed_params(n, re, False))
    warm = sg.Solver(_mixed_params(n, re, True))
    sg.attach_warm_start(warm, net, level)
    teacher = sg.Solver(teacher_params(n, re))
    out, done = [], 0
    H = run.p.H
    for snap in snapshots:
        run.advance(snap - done)
        done = snap
        sysa = {name: run.get(name) for name in P_SYSTEM}
        res = []
        for s in (teacher, zero, warm):
            for name in P_SYSTEM:
                s.set(name, sysa[name])
            s.solve_p()
            res.append(s.get("pc")[H:H + n, H + 1:H + 1 + n])
        conv = res[0] - res[0].mean()
        e = [float(np.linalg.norm(r - r.mean() - conv) / np.linalg.norm(conv)) for r in res[1:]]
        out.append((snap, e[0], e[1]))
    acc, rej = sg.warm_stats(warm)
    for s in (run, zero, warm, teacher):
        s.close()
    return out, (acc, rej)


def ab_to_tolerance(net, n, re, level=-1, cap=10000000, log=print):
    """Outer iterations and seconds to the tolerance, mixed vs learned + mixed."""
    res = {}
    for form in ("mixed", "learned"):
        s = sg.Solver(_mixed_params(n, re, form == "learned", check_every=500))
        if form == "learned":
            sg.attach_warm_start(s, net, level)
        s.advance(3)
        s.autotune(reps=5)  # launch shapes only: results are unchanged
        knobs = sg.tuning_of(s)
        s.close()
        s = sg.Solver(_mixed_params(n, re, form == "learned", check_every=500))
        sg.set_tunings(s, knobs)
        if form == "learned":
            sg.attach_warm_start(s, net, level)
        s.advance(1)
        t0 = time.perf_counter()
        total, conv = 1, False
        while not conv and total < cap:
            k, lg, conv = s.advance(5000)
            total += k
            if k == 0:
                break
        t = time.perf_counter() - t0
        r = {"iterations": total, "seconds": t, "converged": bool(conv),
             "ms_per_iter": 1e3 * t / max(total - 1, 1)}
        if form == "learned":
            r["guesses_kept"], r["guesses_discarded"] = sg.warm_stats(s)
        res[form] = r
        log(f"  {n}^2 Re {re:g} {form}: {r}")
        s.close()
    return res
