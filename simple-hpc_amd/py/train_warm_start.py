"""Trains the learned p' warm start and writes its weights file (warmstart.save format).

Data: batched runs of `--cavities` square cavities (N cycling over 128, 192, 256, 384, 512; Re
log-uniform in [100, 3200]; numpy PCG64 seed 1), each sampled at `--snapshots` outer-iteration
counts spread geometrically from 2 to `--last`; every sample is one (features, converged p')
pair of the cavity's p' system on multigrid level `--level` (warmstart.generate), its target
scaled to unit RMS (warmstart.normalise: the solver rescales the guess by the energy-minimising
factor, so only its shape matters). Training runs on the GPU with the library's bfloat16 MFMA
kernels, each minibatch one random image of its pairs under the 8 symmetries of the square (exact
pairs of the image systems). A held-out set from 8 other cavities (seed 2) gives the reported
errors: the relative RMS error and the shape error after the best scaling.

Usage: python train_warm_start.py OUT.bin [--cavities 128] [--snapshots 32] [--last 8192]
                                  [--steps 12000] [--batch 16] [--lr 1e-3] [--level 2]
Prints one JSON line with the data, timing and held-out figures.
"""
import argparse
import json
import sys
import time

import numpy as np

import simple_gpu as sg
import warmstart as ws


def log(msg):
    print(f"train_warm_start: {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--cavities", type=int, default=128)
    ap.add_argument("--snapshots", type=int, default=32)
    ap.add_argument("--last", type=int, default=8192)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--level", type=int, default=2)
    a = ap.parse_args(argv)
    snaps = sorted(set(int(round(x)) for x in np.geomspace(2, a.last, a.snapshots)))
    cases = ws.training_cavities(n=a.cavities, seed=1)
    t0 = time.perf_counter()
    data = ws.generate(cases, snaps, level=a.level, log=log)
    held = ws.generate(ws.training_cavities(n=8, seed=2), snaps[::4], level=a.level, log=log)
    t_data = time.perf_counter() - t0
    data, held = ws.normalise(data), ws.normalise(held)
    pairs = sum(len(v[0]) for v in data.values())
    log(f"{pairs} pairs ({8 * pairs} with the square's symmetries) in {t_data:.1f} s")
    net = sg.UNet(ws.init_params(0))
    t0 = time.perf_counter()
    hist = ws.train(net, data, steps=a.steps, batch=a.batch, lr=a.lr, log=log, every=1000,
                    augment=True)
    t_train = time.perf_counter() - t0
    mse, mse0 = ws.evaluate(net, held)
    shape = ws.shape_error(net, held)
    ws.save(a.out, net.get_params())
    out = {"pairs": pairs, "pairs_with_symmetries": 8 * pairs, "cavities": a.cavities,
           "snapshots": snaps, "level": a.level, "coarse_shapes": sorted(list(k) for k in data),
           "data_seconds": t_data, "train_steps": a.steps, "batch": a.batch,
           "train_seconds": t_train,
           "loss_first_last": [float(np.mean(hist[:100])), float(np.mean(hist[-100:]))],
           "heldout_pairs": sum(len(v[0]) for v in held.values()),
           "heldout_mse": mse, "heldout_mse_zero": mse0,
           "heldout_rel_rms": float(np.sqrt(mse / mse0)), "heldout_shape_error": shape,
           "weights": a.out}
    print(json.dumps(out), flush=True)
    net.close()


if __name__ == "__main__":
    main()
