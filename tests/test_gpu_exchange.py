"""GPU parity of the exchange build H2ERI_build_exchange.

* a propane molecule built here: GPU K vs the oracle's K on the same
  representation (<= 1e-12; the device adds with atomics, so only the last
  bits depend on the order);
* the reference's fixture K_alk26 (tests/K_alk26_ccpvdz.bin of H2P-ERI):
  GPU K within 5e-6 (the oracle's own distance to the fixture is 1.37e-6 at
  reltol 1e-6), and equal to the oracle;
* device pointers, and repeated calls;
* the density screening of the reference (dlist, |D| <= 1e-10 blocks drop
  out) against the oracle's, and the exchange build as a HIP graph.
"""
import os

import numpy as np
import pytest

import oracle_bind
from conftest import GOLDEN, relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def propane(tmp_path_factory):
    import h2eri
    from h2eri import molecules as M
    rng = M.rng_for(200)
    atoms = M.alkane(3, rng)
    D = M.density(atoms, 13, rng)
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(text=M.mol_text(atoms))
    p = str(tmp_path_factory.mktemp("c3") / "c3.rep")
    h.save(p)
    return h, p, D


def test_exchange_matches_oracle(propane):
    h, p, D = propane
    K = h.build_exchange(D)
    Ko = oracle_bind.Oracle(p).exchange(D)
    assert relerr(K, Ko) < 1e-12
    K2 = h.build_exchange(D)
    assert relerr(K2, K) < 1e-14


def test_exchange_device_pointers(propane):
    import torch
    h, p, D = propane
    Dd = torch.from_numpy(D).to("cuda:0")
    Kd = torch.empty_like(Dd)
    torch.cuda.synchronize()
    h.build_exchange_dev(Dd.data_ptr(), Kd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert relerr(Kd.cpu().numpy(), oracle_bind.Oracle(p).exchange(D)) < 1e-12


def test_exchange_alk26_fixture(tmp_path):
    import h2eri
    h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
    h.build_from_mol(os.path.join(GOLDEN, "alk26_ccpvdz.mol"))
    p = str(tmp_path / "alk26.rep")
    h.save(p)
    nbf = h.nbf
    D = np.fromfile(os.path.join(GOLDEN, "D_alk26_ccpvdz.bin")).reshape(nbf, nbf)
    Kref = np.fromfile(os.path.join(GOLDEN, "K_alk26_ccpvdz.bin")).reshape(nbf, nbf)
    K = h.build_exchange(D)
    assert relerr(K, Kref) < 5e-6
    assert relerr(K, oracle_bind.Oracle(p).exchange(D)) < 1e-12


def test_exchange_density_screening(propane):
    """A density whose shell blocks sit just above or below the screening
    threshold 1e-10 (the reference's dlist): the device K, from host and
    from device pointers, equals the oracle's screened K and differs from the
    K of the same density with every block kept."""
    import torch
    from test_oracle_exchange import np_dlist, screened_test_density
    h, p, D = propane
    o = oracle_bind.Oracle(p)
    bf = h.tree()["shell_bf_sidx"]
    Dx = screened_test_density(D, bf, 5)
    Ko = o.exchange(Dx)
    K = h.build_exchange(Dx)   # also fills the handle's dlist, as the reference's build does
    assert relerr(K, Ko) < 1e-12
    keep = np_dlist(Dx, bf)
    ns = h.s.nshell
    sidx = np.ctypeslib.as_array(h.s.dlist_sidx, shape=(ns + 1,)).copy()
    assert int(sidx[-1]) == int(keep.sum()) and np.array_equal(np.diff(sidx), keep.sum(axis=1))
    Dd = torch.from_numpy(Dx).to("cuda:0")
    Kd = torch.empty_like(Dd)
    h.build_exchange_dev(Dd.data_ptr(), Kd.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert relerr(Kd.cpu().numpy(), Ko) < 1e-12
    # every block kept: a clearly different K (the screening is what matches)
    Dall = np.where(np.abs(Dx) > 0, Dx * 1e3, Dx)   # all blocks above the threshold
    assert relerr(h.build_exchange(Dall) * 1e-3, Ko) > 1e-2


def test_exchange_scratch_chunks(propane):
    """The block terms run in chunks whose intermediate G fits a bound
    (here 64 KB, so nearly every term is a chunk of its own; even chunks on
    the caller's stream, odd ones on a second stream in a second region):
    K equals the one-chunk build to 1e-14 (FP64 atomics order the adds
    differently) and the oracle to 1e-12, from host and device pointers and
    as a graph replayed twice, and the resident scratch shrinks."""
    import torch
    h, p, D = propane
    h.set_exchange_scratch(1 << 40)
    K1 = h.build_exchange(D)
    big = h.exchange_stats()["scratch_doubles"]
    try:
        h.set_exchange_scratch(65536)
        Kc = h.build_exchange(D)
        small = h.exchange_stats()["scratch_doubles"]
        Dd = torch.from_numpy(D).to("cuda:0")
        Kd = torch.empty_like(Dd)
        s = torch.cuda.current_stream().cuda_stream
        h.build_exchange_dev(Dd.data_ptr(), Kd.data_ptr(), s)
        torch.cuda.synchronize()
        assert relerr(Kd.cpu().numpy(), Kc) < 1e-14
        Kg = torch.empty_like(Dd)
        h.graph_capture("exchange", 1, Dd.data_ptr(), Kg.data_ptr())
        for _ in range(2):
            Kg.zero_()
            h.graph_replay(s)
            torch.cuda.synchronize()
            assert relerr(Kg.cpu().numpy(), Kc) < 1e-14
    finally:
        h.set_exchange_scratch(2 << 30)   # the default
    assert relerr(Kc, K1) < 1e-14
    assert relerr(Kc, oracle_bind.Oracle(p).exchange(D)) < 1e-12
    assert small < big


@pytest.mark.parametrize("name", ["propane", "water8"])
def test_exchange_graph_replay(propane, tmp_path_factory, name):
    """The exchange build captured in a HIP graph (kind 2) and replayed three
    times with new densities and eager work between the replays: every replay
    equals the eager device call to 1e-14 (FP64 atomics order the adds
    differently) and the oracle to 1e-12."""
    import torch
    import h2eri
    from h2eri import molecules as M
    if name == "propane":
        h, p, D = propane
        atoms, nocc = M.alkane(3, M.rng_for(200)), 13
    else:
        rng = M.rng_for(400)
        atoms = M.water_cluster(8, rng)
        D = M.density(atoms, 40, rng)
        nocc = 40
        h = h2eri.H2ERI(1e-10, 1e-10, 1e-6)
        h.build_from_mol(text=M.mol_text(atoms))
        p = str(tmp_path_factory.mktemp("w8k") / "w8.rep")
        h.save(p)
    o = oracle_bind.Oracle(p)
    Dd = torch.from_numpy(D).to("cuda:0")
    Kg = torch.empty_like(Dd)
    Ke = torch.empty_like(Dd)
    s = torch.cuda.current_stream().cuda_stream
    h.graph_capture("exchange", 1, Dd.data_ptr(), Kg.data_ptr())
    rng = np.random.default_rng(11)
    for k in range(3):
        Dk = M.density(atoms, nocc, rng)
        Dd.copy_(torch.from_numpy(Dk))
        Kg.zero_()
        h.graph_replay(s)
        h.build_exchange_dev(Dd.data_ptr(), Ke.data_ptr(), s)
        torch.cuda.synchronize()
        Kgh = Kg.cpu().numpy()
        assert relerr(Kgh, Ke.cpu().numpy()) < 1e-14
        assert relerr(Kgh, o.exchange(Dk)) < 1e-12
    # freeing the exchange data releases its graph: the replay is refused
    h.exchange_workbuf_free()
    with pytest.raises(h2eri.H2ERIError):
        h.graph_replay(s)
