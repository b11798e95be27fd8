"""GPU parity at the full size of the benchmark workloads.

Each workload is built once per session with the benchmark's own seeds and
sizes (h2eri.molecules), saved, and read by the oracle:

* C30H62, cc-pVDZ, reltol 1e-6, one density (nocc 121): J in FP64 against
  the oracle <= 1e-13, with both executors; the FP32-stored J within the
  precision gate 1e-7 of the FP64 J;
* C100H202, reltol 1e-8 (the largest single-GPU representation): J in FP64
  against the oracle <= 1e-13 and the FP32-stored J within 1e-7;
* the batching workload on that representation: 16 seeded densities (nocc
  401) in one batched FP32-stored call, each equal to its single-density
  call <= 1e-13 and two of them against the oracle's FP64 J within the
  precision gate; and the eight alkanes C20..C160 (reltol 1e-6,
  FP32-stored) as one molecule group, each against its oracle FP64 matvec
  within the precision gate;
* (H2O)64, reltol 1e-6: J in FP64 against the oracle <= 1e-13, FP32-stored
  J within 1e-7, and K (nocc 320) against the oracle <= 1e-12.

Together these take minutes (host construction of the representations,
oracle products on the host cores), each test well under two; the small
analogues in the other files cover the same code paths in seconds.
"""
import time

import numpy as np
import pytest

import oracle_bind
from conftest import relerr

pytestmark = pytest.mark.gpu

EPS_PREC = 1e-7   # precision gate: FP32-stored U/B against FP64 on the same representation


def _log(msg):
    print(f"[workloads] {msg}", flush=True)


def _build(tmp_path_factory, name, atoms, reltol):
    import h2eri
    from h2eri import molecules as M
    t0 = time.perf_counter()
    h = h2eri.H2ERI(1e-10, 1e-10, reltol)
    h.build_from_mol(text=M.mol_text(atoms))
    p = str(tmp_path_factory.mktemp(name) / f"{name}.rep")
    h.save(p)
    _log(f"{name}: {h.nbf} basis functions, N={h.n}, built in {time.perf_counter() - t0:.1f} s")
    return h, p


def _workload(tmp_path_factory, config):
    from h2eri import molecules as M
    atoms, _, D, reltol = M.workload(config)
    h, p = _build(tmp_path_factory, M.workload_name(config).replace("(", "").replace(")", ""), atoms, reltol)
    h.atoms = atoms
    return h, p, D


@pytest.fixture(scope="module")
def c30(tmp_path_factory):
    return _workload(tmp_path_factory, 1)


@pytest.fixture(scope="module")
def c100(tmp_path_factory):
    return _workload(tmp_path_factory, 2)


@pytest.fixture(scope="module")
def water64(tmp_path_factory):
    return _workload(tmp_path_factory, 4)


def _j_checks(h, path, D, executors=(0,)):
    o = oracle_bind.Oracle(path)
    Jo = o.coulomb(D)
    o.close()
    J64 = None
    for ex in executors:
        h.set_executor(ex)
        try:
            J = h.build_coulomb(D)
        finally:
            h.set_executor(0)
        e = relerr(J, Jo)
        _log(f"N={h.n} executor {ex}: J FP64 vs oracle {e:.2e}")
        assert e < 1e-13
        J64 = J
    h.set_precision(32)
    try:
        J32 = h.build_coulomb(D)
    finally:
        h.set_precision(64)
    eps = relerr(J32, J64)
    _log(f"N={h.n}: J FP32-stored vs FP64 {eps:.2e}")
    assert 0.0 < eps < EPS_PREC
    return Jo


def test_c30h62_j(c30):
    h, p, D = c30
    _j_checks(h, p, D, executors=(0, 1))


def test_c100h202_j(c100):
    h, p, D = c100
    _j_checks(h, p, D, executors=(0, 1))


def test_c100h202_sixteen_densities(c100):
    """The batching workload's 16 densities (seeds of the benchmark, nocc 401)
    in one FP32-stored batched call: two passes of 8 over the matrix data."""
    from h2eri import molecules as M
    h, p, _ = c100
    Ds = M.batch_densities(h.atoms, 401, 16)
    h.set_precision(32)
    try:
        Js = h.build_coulomb_multi(Ds)
        singles = [h.build_coulomb(Ds[k]) for k in (0, 7, 8, 15)]
    finally:
        h.set_precision(64)
    for k, J1 in zip((0, 7, 8, 15), singles):
        e = relerr(Js[k], J1)
        _log(f"density {k}: batched vs single {e:.2e}")
        assert e < 1e-13
    o = oracle_bind.Oracle(p)
    for k in (0, 15):
        e = relerr(Js[k], o.coulomb(Ds[k]))
        _log(f"density {k}: FP32-stored batched vs oracle FP64 {e:.2e}")
        assert e < EPS_PREC
    o.close()


ALKANES = list(range(20, 161, 20))
_group = {}   # C count -> (handle, representation path, x): built one test at a time


@pytest.mark.parametrize("n", ALKANES)
def test_alkane_group_member(tmp_path_factory, n):
    """Builds one molecule of the batching workload's group (one test each,
    so that no single test runs for minutes) and checks its FP32-stored
    matvec against the oracle's FP64 matvec within the precision gate."""
    from h2eri import molecules as M
    rng = M.rng_for(300 + n)
    atoms = M.alkane(n, rng)
    h, p = _build(tmp_path_factory, f"C{n}", atoms, 1e-6)
    h.set_precision(32)
    x = np.random.default_rng(n).standard_normal(h.n)
    o = oracle_bind.Oracle(p)
    e = relerr(h.matvec(x), o.matvec(x))
    o.close()
    assert e < EPS_PREC
    _group[n] = (h, p, x)


def test_alkane_group_c20_to_c160():
    """The eight alkanes C20H42..C160H322 (reltol 1e-6, FP32-stored) as one
    molecule group, one persistent launch; each molecule's y against its
    oracle FP64 matvec within the precision gate, and bitwise equal to the
    FP32-stored persistent matvec of the molecule alone."""
    import torch
    import h2eri
    assert sorted(_group) == ALKANES, "the member tests build the molecules"
    hs = [_group[n][0] for n in ALKANES]
    X = [torch.from_numpy(_group[n][2]).to("cuda:0") for n in ALKANES]
    Y = [torch.zeros_like(x) for x in X]
    g = h2eri.Group(hs, [x.data_ptr() for x in X], [y.data_ptr() for y in Y])
    try:
        g.matvec(torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        assert g.status() == 0
        for k, n in enumerate(ALKANES):
            h, p, x = _group[n]
            o = oracle_bind.Oracle(p)
            e = relerr(Y[k].cpu().numpy(), o.matvec(x))
            o.close()
            _log(f"C{n}: group (FP32-stored) vs oracle {e:.2e}")
            assert e < EPS_PREC
            h.set_executor(1)
            assert np.array_equal(Y[k].cpu().numpy(), h.matvec(x))
            h.set_executor(0)
    finally:
        g.close()
        _group.clear()


def test_water64_j_and_k(water64):
    h, p, D = water64
    _j_checks(h, p, D)
    t0 = time.perf_counter()
    K = h.build_exchange(D)
    _log(f"(H2O)64: K on the device in {time.perf_counter() - t0:.1f} s (plan, upload and basis products included)")
    h.exchange_workbuf_free()
    o = oracle_bind.Oracle(p)
    t0 = time.perf_counter()
    Ko = o.exchange(D)
    o.close()
    e = relerr(K, Ko)
    _log(f"(H2O)64: K vs oracle {e:.2e} (oracle {time.perf_counter() - t0:.1f} s)")
    assert e < 1e-12
