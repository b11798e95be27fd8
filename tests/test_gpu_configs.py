"""GPU parity on small analogues of the benchmark configurations.

* 3D globular case with exchange: a cluster of 8 water molecules (O and H
  cc-pVDZ blocks): J in FP64 and FP32-stored form, and K, against the oracle.
* tight tolerance: a C8H18 chain at reltol 1e-8 (deeper ranks): J in FP64
  against the oracle and the FP32-stored J error.
* many densities: 16 damped densities on one representation in one batched
  call, each against the oracle.
* uneven molecules: alkanes of 3, 5 and 9 carbons, each handle with its own
  density, J against the oracle.
"""
import numpy as np
import pytest

import oracle_bind
from conftest import relerr

pytestmark = pytest.mark.gpu


def build(tmp_path_factory, name, atoms, reltol):
    import h2eri
    from h2eri import molecules as M
    h = h2eri.H2ERI(1e-10, 1e-10, reltol)
    h.build_from_mol(text=M.mol_text(atoms))
    p = str(tmp_path_factory.mktemp(name) / f"{name}.rep")
    h.save(p)
    return h, p


@pytest.fixture(scope="module")
def water8(tmp_path_factory):
    from h2eri import molecules as M
    rng = M.rng_for(400)
    atoms = M.water_cluster(8, rng)
    D = M.density(atoms, 40, rng)
    h, p = build(tmp_path_factory, "w8", atoms, 1e-6)
    return h, p, D


def test_water_cluster_j_and_k(water8):
    h, p, D = water8
    o = oracle_bind.Oracle(p)
    assert h.s.n_B > 0
    J = h.build_coulomb(D)
    assert relerr(J, o.coulomb(D)) < 1e-13
    h.set_precision(32)
    try:
        J32 = h.build_coulomb(D)
    finally:
        h.set_precision(64)
    assert relerr(J32, J) < 1e-7   # precision gate of the FP32-stored form
    K = h.build_exchange(D)
    assert relerr(K, o.exchange(D)) < 1e-12


def test_tight_tolerance_chain(tmp_path_factory):
    from h2eri import molecules as M
    rng = M.rng_for(401)
    atoms = M.alkane(8, rng)
    D = M.density(atoms, 33, rng)
    h, p = build(tmp_path_factory, "c8t", atoms, 1e-8)
    o = oracle_bind.Oracle(p)
    J = h.build_coulomb(D)
    Jo = o.coulomb(D)
    assert relerr(J, Jo) < 1e-13
    h.set_precision(32)
    try:
        assert relerr(h.build_coulomb(D), Jo) < 1e-7   # precision gate of the FP32-stored form
    finally:
        h.set_precision(64)


def test_sixteen_densities(water8):
    from h2eri import molecules as M
    h, p, D0 = water8
    o = oracle_bind.Oracle(p)
    rng = np.random.default_rng(5)
    atoms = M.water_cluster(8, M.rng_for(400))
    Ds = np.array([M.density(atoms, 40, rng) for _ in range(16)])
    Js = h.build_coulomb_multi(Ds)
    for k in range(16):
        assert relerr(Js[k], o.coulomb(Ds[k])) < 1e-13


def test_uneven_molecules(tmp_path_factory):
    from h2eri import molecules as M
    for n in (3, 5, 9):
        rng = M.rng_for(500 + n)
        atoms = M.alkane(n, rng)
        D = M.density(atoms, 4 * n + 1, rng)
        h, p = build(tmp_path_factory, f"c{n}", atoms, 1e-6)
        assert relerr(h.build_coulomb(D), oracle_bind.Oracle(p).coulomb(D)) < 1e-13


def test_molecule_group(tmp_path_factory):
    """Uneven molecules in one persistent launch (H2ERI_group_*): each equals
    the oracle, and bitwise the molecule's own persistent matvec (same plan,
    same sums)."""
    import torch
    import h2eri
    from h2eri import molecules as M
    hs, paths, xs = [], [], []
    for n in (4, 7, 11):
        rng = M.rng_for(600 + n)
        atoms = M.alkane(n, rng)
        h, p = build(tmp_path_factory, f"g{n}", atoms, 1e-6)
        hs.append(h)
        paths.append(p)
        xs.append(np.random.default_rng(n).standard_normal(h.n))
    X = [torch.from_numpy(x).to("cuda:0") for x in xs]
    Y = [torch.zeros_like(x) for x in X]
    g = h2eri.Group(hs, [x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    for rep in range(2):
        for y in Y:
            y.zero_()
        g.matvec(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert g.status() == 0
        for k in range(3):
            assert relerr(Y[k].cpu().numpy(), oracle_bind.Oracle(paths[k]).matvec(xs[k])) < 1e-13
            hs[k].set_executor(1)
            assert np.array_equal(Y[k].cpu().numpy(), hs[k].matvec(xs[k]))
            hs[k].set_executor(0)
    # mode 1: each molecule's own pass (its handle's configuration, here the
    # staged executor): bitwise its own matvec; the tuner keeps either mode
    g.set_mode(1)
    for y in Y:
        y.zero_()
    g.matvec(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert g.status() == 0
    for k in range(3):
        assert np.array_equal(Y[k].cpu().numpy(), hs[k].matvec(xs[k]))
    res = g.autotune(reps=2)
    assert res["mode"] in (0, 1) and res["mode0_ms"] > 0 and res["mode1_ms"] > 0
    for y in Y:
        y.zero_()
    g.matvec(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for k in range(3):
        assert relerr(Y[k].cpu().numpy(), oracle_bind.Oracle(paths[k]).matvec(xs[k])) < 1e-13
    g.close()


def test_autotune_keeps_results(tmp_path):
    from conftest import make_synthetic
    h = make_synthetic(tmp_path / "a.rep", 2000, 21, leaf=60, max_rank=30, lr_frac=0.4, diag_lr_frac=0.3)
    res = h.autotune(nv=1, reps=3)
    assert res["ms"] > 0 and res["tile_bytes"] in (65536, 131072, 262144, 524288)
    assert res["schedule"] in (0, 2, 3, 4, 6, 10) and res["near_field_grid"] >= 0
    assert res["executor"] in (0, 1) and h.executor() == res["executor"] and res["wave_tile_bytes"] in (0, 32768)
    x = np.random.default_rng(2).standard_normal(h.n)
    o = oracle_bind.Oracle(str(tmp_path / "a.rep"))
    assert relerr(h.matvec(x), o.matvec(x)) < 1e-13
    # every batch width keeps a configuration of its own: tuning 8 vectors per
    # pass leaves the single-vector configuration alone
    res8 = h.autotune(nv=8, reps=3)
    assert res8["executor"] in (0, 1) and h.executor() == res["executor"]
    X = np.random.default_rng(3).standard_normal((9, h.n))   # one pass of 8, one of 1
    Y = h.matvec_multi(X)
    for v in range(9):
        assert relerr(Y[v], o.matvec(X[v])) < 1e-13
