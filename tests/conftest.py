import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "h2p-eri_amd"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def synthetic_points(n, seed, shape="chain"):
    """Seeded shell-pair centres and sizes for chemistry-free trees."""
    rng = np.random.default_rng(seed)
    pts = np.zeros((3, n))
    if shape == "chain":
        pts[0] = rng.uniform(0, 0.06 * n, n)
        pts[1] = rng.uniform(0, 3, n)
        pts[2] = rng.uniform(0, 3, n)
    else:
        pts[:] = rng.uniform(0, 2.0 * n ** (1 / 3), (3, n))
    nbfp = rng.choice([1, 3, 6, 9, 18, 36], n).astype(np.int32)
    return pts, nbfp


def make_synthetic(path, n, seed, leaf=30, max_rank=12, lr_frac=0.4, diag_lr_frac=0.0, shape="chain"):
    import h2eri
    pts, nbfp = synthetic_points(n, seed, shape)
    h = h2eri.H2ERI().partition_synthetic(pts, nbfp, leaf)
    h.build_synthetic(seed=seed, max_rank=max_rank, lr_frac=lr_frac, diag_lr_frac=diag_lr_frac)
    h.save(str(path))
    return h


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))
