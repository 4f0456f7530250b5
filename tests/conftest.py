"""Shared test setup: the `gpu` marker, import paths and small helpers.

CPU tests (`-m "not gpu"`) check the oracle against the reference's fixtures, invariants and
(when it is built here) the reference itself, plus the host code and the C ABI's symbol table.
GPU tests (`-m gpu`) check the HIP backend, through the C ABI, against the oracle.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "simple-hpc_amd", "py"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP backend)")


def golden_cfg(name):
    with open(os.path.join(GOLDEN, "snapshots", name + ".cfg")) as f:
        return f.read()


def with_keys(cfg, **kv):
    """Config text with keys replaced or added (like run_bench.py --set KEY=VALUE)."""
    lines = []
    for line in cfg.splitlines():
        key = line.split("#")[0].split("=")[0].strip()
        if key in kv:
            continue
        lines.append(line)
    lines += [f"{k} = {v}" for k, v in kv.items()]
    return "\n".join(lines) + "\n"


def load_csv(path):
    return np.loadtxt(path, delimiter=",", ndmin=2)


def gpu_available():
    try:
        import simple_gpu
        return simple_gpu.device_count() > 0
    except Exception:
        return False


def bits_equal(a, b):
    """Bit-for-bit equality of two float64 arrays (NaN payloads and signed zeros included)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def first_diff(a, b):
    d = np.argwhere(a.view(np.uint64) != b.view(np.uint64))
    if len(d) == 0:
        return "identical"
    i, j = d[0]
    rel = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    return f"first diff at ({i},{j}): {a[i, j]!r} vs {b[i, j]!r}; max rel {rel:.3e}; n={len(d)}"
