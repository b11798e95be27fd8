"""Pins of the CPU oracle's matvec (CPU only).

The sweep-based restatement orc_matvec must equal the dense brute-force
assembly orc_dense_matrix (an independent formulation of the same H2 matrix)
on trees that exercise all three coupling cases and low-rank B/D blocks.
"""
import numpy as np
import pytest

import oracle_bind
from conftest import make_synthetic, relerr


@pytest.mark.parametrize("seed,n,shape,diag_lr", [(1, 500, "chain", 0.0), (2, 700, "chain", 0.5),
                                                   (3, 400, "cube", 0.3), (4, 60, "chain", 0.0)])
def test_oracle_matches_dense(tmp_path, seed, n, shape, diag_lr):
    h = make_synthetic(tmp_path / "r.rep", n, seed, leaf=20, max_rank=10, lr_frac=0.5, diag_lr_frac=diag_lr,
                       shape=shape)
    tr = h.tree()
    o = oracle_bind.Oracle(str(tmp_path / "r.rep"))
    A = o.dense()
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3, o.n))
    for x in X:
        assert relerr(o.matvec(x), A @ x) < 1e-13
    if n >= 400 and shape == "chain":
        lv = tr["node_level"][tr["r_adm_pairs"]]
        assert (lv[:, 0] == lv[:, 1]).any() and (lv[:, 0] > lv[:, 1]).any() and (lv[:, 0] < lv[:, 1]).any()


def test_oracle_linear_and_symmetric(tmp_path):
    # dense diagonal blocks are symmetric and every off-diagonal block enters
    # with its transpose, so A_H2 is symmetric
    make_synthetic(tmp_path / "r.rep", 600, 7, leaf=25, max_rank=14, lr_frac=0.5)
    o = oracle_bind.Oracle(str(tmp_path / "r.rep"))
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal(o.n), rng.standard_normal(o.n)
    Au, Av = o.matvec(u), o.matvec(v)
    assert abs(u @ Av - v @ Au) <= 1e-12 * np.linalg.norm(Au) * np.linalg.norm(v)
    assert relerr(o.matvec(2.0 * u - 3.0 * v), 2.0 * Au - 3.0 * Av) < 1e-13


def test_oracle_dense_catches_dropped_transpose(tmp_path):
    # planted mistake on the CPU: dropping the transposed half of the
    # inadmissible blocks changes the product, so the check is sensitive
    make_synthetic(tmp_path / "r.rep", 500, 8, leaf=20, max_rank=10, lr_frac=0.0)
    o = oracle_bind.Oracle(str(tmp_path / "r.rep"))
    A = o.dense()
    Lt = np.tril(A, -1)
    assert np.linalg.norm(Lt) > 0 and relerr(A - Lt, A) > 1e-3
